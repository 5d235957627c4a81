// efvm_stage_f32.hip -- FP32 production stage kernel (contraction allowed).
//
// Same wave-strip decomposition and boundary semantics as the FP64 parity
// kernel (efvm_stage_f64.hip); the numerics are the 2-D specialised schemes of
// efvm_physics.h (fast::).  The z faces of the one-cell-thick mesh are dropped:
// their two contributions cancel exactly in exact arithmetic.  The divergence
// is (F_w - F_e)*A_x/V + (G_s - G_n)*A_y/V, then viscous, body force, solid
// zero and the RK combine as FMAs.  Stage 3 fuses the CFL scan: lambda of the
// new state per cell (FP32), max-reduced in FP64 across the wave and the grid.
#include "efvm_stage.h"

namespace efvm {

using F4 = fast::C4<float>;
using FP = fast::Prep<float>;

__device__ __forceinline__ F4 shfl_up4(const F4& v) { return {shfl_up1(v.r), shfl_up1(v.ru), shfl_up1(v.rv), shfl_up1(v.re)}; }
__device__ __forceinline__ F4 shfl_dn4(const F4& v) { return {shfl_dn1(v.r), shfl_dn1(v.ru), shfl_dn1(v.rv), shfl_dn1(v.re)}; }
__device__ __forceinline__ FP shfl_upP(const FP& v) {
    return {shfl_up1(v.r), shfl_up1(v.ru), shfl_up1(v.rv), shfl_up1(v.re), shfl_up1(v.u), shfl_up1(v.v), shfl_up1(v.p), shfl_up1(v.a)};
}
__device__ __forceinline__ F4 as4(const Cell<float>& c) { return {c.r, c.ru, c.rv, c.re}; }

template <int LIM>
__device__ __forceinline__ void recon_pair(const F4& um, const F4& uc, const F4& up, bool rec_lo, bool rec_hi, F4& lo, F4& hi) {
    F4 s;
    s.r = fast::limit<LIM>(up.r - uc.r, uc.r - um.r);
    s.ru = fast::limit<LIM>(up.ru - uc.ru, uc.ru - um.ru);
    s.rv = fast::limit<LIM>(up.rv - uc.rv, uc.rv - um.rv);
    s.re = fast::limit<LIM>(up.re - uc.re, uc.re - um.re);
    lo = uc; hi = uc;
    if (rec_lo) { lo.r -= 0.5f * s.r; lo.ru -= 0.5f * s.ru; lo.rv -= 0.5f * s.rv; lo.re -= 0.5f * s.re; }
    if (rec_hi) { hi.r += 0.5f * s.r; hi.ru += 0.5f * s.ru; hi.rv += 0.5f * s.rv; hi.re += 0.5f * s.re; }
}

// One wave processes strip `s` of the case described by `a`; the planes and the
// stage/parity are passed separately so the persistent batched kernel can reuse it.
// The wave owning strip 0 does the per-case clock bookkeeping.
template <int FLUX, int LIM>   // LIM = LIM_NONE: first order
__device__ __forceinline__ void stage_f32_strip(const StageArgs& a, const int s, const float* const* U,
                                                const float* const* U0, float* const* O, const int stage,
                                                const int parity) {
    constexpr bool MUSCL = LIM != LIM_NONE;
    const int lane = lane_id();
    const int nx = a.nx, ny = a.ny;
    const int sxi = s % a.n_strips_x, syi = s / a.n_strips_x;
    const int c = sxi * STRIP_OUT - 2 + lane;
    const int j0 = syi * a.strip_h;
    const int j1 = min(j0 + a.strip_h, ny);
    const float g = (float)a.gamma, gm1 = (float)(a.gamma - 1.0);
    const double dtd = a.fixed_dt > 0.0 ? a.fixed_dt : a.cfl / a.clock->lam[parity];
    const float dt = (float)dtd;
    const bool fuse = a.fuse_cfl && (a.euler || stage == 2);
    if (s == 0 && lane == 0 && stage == 0) { a.clock->dt = dtd; a.clock->lam[parity ^ 1] = 0.0; }
    float ca, cb;   // out = ca*u0 + cb*(u + dt*rhs)
    if (a.euler || stage == 0) { ca = 0.f; cb = 1.f; }
    else if (stage == 1) { ca = 0.75f; cb = 0.25f; }
    else { ca = 1.f / 3.f; cb = 2.f / 3.f; }
    const float ix = (float)a.inv_x, iy = (float)a.inv_y;

    const bool col_ghost = (c < 0 || c >= nx);
    const bool seam_rec = a.seam_mode == 1;
    const bool xW_rec = MUSCL && (col_ghost ? (a.per_x && seam_rec) : !(c == 0 && !a.per_x));
    const bool xE_rec = MUSCL && (col_ghost ? (a.per_x && seam_rec) : !(c == nx - 1 && !a.per_x));
    const bool out_lane = lane >= 2 && lane < WAVE - 2 && c < nx;
    const bool visc = a.viscous != 0;
    const bool ibm = a.solid != nullptr;

    F4 um = as4(load_cell(a, U, c, j0 - 2));
    F4 uc = as4(load_cell(a, U, c, j0 - 1));
    F4 up = as4(load_cell(a, U, c, j0));
    int sm = 0, scl = 0, spl = 0;
    auto solid_at = [&](int cc, int rr) -> int {
        if (!ibm || cc < 0 || cc >= nx || rr < 0 || rr >= ny) return 0;
        return a.solid[(long)rr * nx + cc];
    };
    if (ibm) { sm = solid_at(c, j0 - 2); scl = solid_at(c, j0 - 1); spl = solid_at(c, j0); }

    FP hi_prev{};
    F4 G_prev{0.f, 0.f, 0.f, 0.f}, dX_prev{0.f, 0.f, 0.f, 0.f};
    float vis_prev[3] = {0.f, 0.f, 0.f};
    float lam_max = 0.f;
    int bad = 0;

    for (int jc = j0 - 1; jc <= j1; ++jc) {
        const bool row_ghost = (jc < 0 || jc >= ny);
        const bool yS_rec = MUSCL && (row_ghost ? (a.per_y && seam_rec) : !(jc == 0 && !a.per_y));
        const bool yN_rec = MUSCL && (row_ghost ? (a.per_y && seam_rec) : !(jc == ny - 1 && !a.per_y));
        FP lo_p, hi_p;
        if (MUSCL) {
            F4 lo, hi;
            recon_pair<LIM>(um, uc, up, yS_rec, yN_rec, lo, hi);
            lo_p = fast::prep(lo, g, gm1);
            hi_p = fast::prep(hi, g, gm1);
        } else {
            lo_p = fast::prep(uc, g, gm1);
            hi_p = lo_p;
        }
        F4 G{0.f, 0.f, 0.f, 0.f};
        if (jc >= j0) G = fast::flux<FLUX, 1>(hi_prev, lo_p, g, gm1);

        F4 dX{0.f, 0.f, 0.f, 0.f};
        float vis[3] = {0.f, 0.f, 0.f};
        if (jc >= j0 && jc < j1) {
            FP W_p, E_p;
            if (MUSCL) {
                F4 ul = shfl_up4(uc), ur = shfl_dn4(uc);
                F4 W, E;
                recon_pair<LIM>(ul, uc, ur, xW_rec, xE_rec, W, E);
                W_p = fast::prep(W, g, gm1);
                E_p = fast::prep(E, g, gm1);
            } else {
                W_p = (jc == j0 - 1) ? lo_p : lo_p;   // first order: both x states are the cell value
                E_p = W_p;
            }
            FP El = shfl_upP(E_p);
            F4 Fw = fast::flux<FLUX, 0>(El, W_p, g, gm1);
            F4 Fe = shfl_dn4(Fw);
            dX = {Fw.r - Fe.r, Fw.ru - Fe.ru, Fw.rv - Fe.rv, Fw.re - Fe.re};
            if (visc) {
                // primitives of the centre row and its x-neighbours
                float ir = fast::rcp(uc.r);
                float pu = uc.ru * ir, pv = uc.rv * ir;
                float pt = (gm1 * (uc.re - 0.5f * (uc.ru * pu + uc.rv * pv))) * ir / (float)a.r_gas;
                float lu = shfl_up1(pu), lv = shfl_up1(pv), lt = shfl_up1(pt);
                float ru_ = shfl_dn1(pu), rv_ = shfl_dn1(pv), rt = shfl_dn1(pt);
                int sl_ = shfl_up1(scl), sr_ = shfl_dn1(scl);
                auto prim_of = [&](const F4& q, float& u, float& v, float& t) {
                    float iq = fast::rcp(q.r);
                    u = q.ru * iq; v = q.rv * iq;
                    t = (gm1 * (q.re - 0.5f * (q.ru * u + q.rv * v))) * iq / (float)a.r_gas;
                };
                float du, dv, dtt, uu, vu, tu;
                prim_of(um, du, dv, dtt);
                prim_of(up, uu, vu, tu);
                if (!scl && !col_ghost) {
                    float mu = a.visc_const ? (float)a.mu : (float)ref::sutherland((double)pt);
                    float kap = mu * (float)(a.gamma * a.r_gas / (a.gamma - 1.0) / 0.72);
                    if (sr_) { ru_ = 0.f; rv_ = 0.f; rt = pt; }
                    if (sl_) { lu = 0.f; lv = 0.f; lt = pt; }
                    if (spl) { uu = 0.f; vu = 0.f; tu = pt; }
                    if (sm) { du = 0.f; dv = 0.f; dtt = pt; }
                    const float yim = (float)a.yw[jc], yip = (float)a.yw[ny + jc], yis = (float)a.yw[2 * ny + jc];
                    const float dxi2 = (float)a.dxi2;
                    float lapu = (ru_ - 2.f * pu + lu) * dxi2 + ((uu - pu) * yip - (pu - du) * yim) * yis;
                    float lapv = (rv_ - 2.f * pv + lv) * dxi2 + ((vu - pv) * yip - (pv - dv) * yim) * yis;
                    float lapt = (rt - 2.f * pt + lt) * dxi2 + ((tu - pt) * yip - (pt - dtt) * yim) * yis;
                    vis[0] = mu * lapu; vis[1] = mu * lapv;
                    vis[2] = pu * vis[0] + pv * vis[1] + kap * lapt;
                }
            }
        }

        if (jc - 1 >= j0 && out_lane) {
            const int j = jc - 1;
            const long q = (long)j * nx + c;
            float rr = dX_prev.r * ix + (G_prev.r - G.r) * iy;
            float rru = dX_prev.ru * ix + (G_prev.ru - G.ru) * iy + vis_prev[0];
            float rrv = dX_prev.rv * ix + (G_prev.rv - G.rv) * iy + vis_prev[1];
            float rre = dX_prev.re * ix + (G_prev.re - G.re) * iy + vis_prev[2];
            if (a.body_force) { float gx = (float)a.gx; rru += um.r * gx; rre += gx * um.ru; }
            if (ibm && a.solid[q]) { rr = 0.f; rru = 0.f; rrv = 0.f; rre = 0.f; }
            float o0 = um.r + dt * rr, o1 = um.ru + dt * rru, o2 = um.rv + dt * rrv, o3 = um.re + dt * rre;
            if (ca != 0.f) {
                o0 = ca * U0[0][q] + cb * o0; o1 = ca * U0[1][q] + cb * o1;
                o2 = ca * U0[2][q] + cb * o2; o3 = ca * U0[3][q] + cb * o3;
            }
            O[0][q] = o0; O[1][q] = o1; O[2][q] = o2; O[3][q] = o3;
            if (fuse) {
                float ir = fast::rcp(o0);
                float vu = fabsf(o1 * ir), vv = fabsf(o2 * ir);
                float p = gm1 * (o3 - 0.5f * (o1 * o1 + o2 * o2) * ir);
                float snd = fast::fsqrt(g * p * ir);
                float l = (vu + snd) * (float)(1.0 / a.dx) + (vv + snd) * (float)(1.0 / a.dy) + snd * (float)(1.0 / a.dz);
                if (a.viscous) {
                    float t = p * ir / (float)a.r_gas;
                    float mu = a.visc_const ? (float)a.mu : (float)ref::sutherland((double)t);
                    l += 2.f * mu * ir * (float)(1.0 / (a.dx * a.dx) + 1.0 / (a.dy * a.dy) + 1.0 / (a.dz * a.dz));
                }
                if (o0 > 0.f && l == l && p > 0.f) lam_max = fmaxf(lam_max, l);
                else bad++;
            }
        }

        hi_prev = hi_p; G_prev = G; dX_prev = dX;
        vis_prev[0] = vis[0]; vis_prev[1] = vis[1]; vis_prev[2] = vis[2];
        um = uc; uc = up;
        if (ibm) { sm = scl; scl = spl; }
        if (jc + 2 <= j1 + 1) {
            up = as4(load_cell(a, U, c, jc + 2));
            if (ibm) spl = solid_at(c, jc + 2);
        }
    }
    if (fuse) {
        double w = wave_max((double)lam_max);
        long long b = wave_sum_ll((long long)bad);
        if (lane == 0) {
            if (w > 0.0) atomic_max_pos_double(&a.clock->lam[parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    if (s == 0 && lane == 0 && (a.euler || stage == 2)) {
        a.clock->time += dtd;
        a.clock->step += 1;
    }
}

template <int FLUX, int LIM>
__global__ __launch_bounds__(256) void stage_f32_kernel(StageArgs a) {
    const int s = strip_index(a, blockDim.x / WAVE);
    if (s >= a.n_strips_x * a.n_strips_y) return;
    const float* U[4] = {(const float*)a.uin[0], (const float*)a.uin[1], (const float*)a.uin[2], (const float*)a.uin[3]};
    const float* U0[4] = {(const float*)a.u0[0], (const float*)a.u0[1], (const float*)a.u0[2], (const float*)a.u0[3]};
    float* O[4] = {(float*)a.uout[0], (float*)a.uout[1], (float*)a.uout[2], (float*)a.uout[3]};
    stage_f32_strip<FLUX, LIM>(a, s, U, U0, O, a.stage, a.parity);
}

// ---------------------------------------------------------------------------
// Persistent batched stage: hundreds of independent cases of uneven size in one
// launch.  A host-built work table lists (case, strip) pairs, largest strips
// first; a fixed grid of waves strides through it.  Every case keeps its own
// DevClock (per-case dt from its own CFL maximum) and its own plane sets
// A, B, C; all cases advance in lock-step stage by stage.
// ---------------------------------------------------------------------------
template <int FLUX, int LIM>
__global__ __launch_bounds__(256) void stage_f32_batch_kernel(const BatchCase* __restrict__ cases,
                                                              const int2* __restrict__ work, int n_work,
                                                              int stage, int parity) {
    const int wpb = blockDim.x / WAVE;
    const int nw = gridDim.x * wpb;
    for (int w = blockIdx.x * wpb + (int)(threadIdx.x / WAVE); w < n_work; w += nw) {
        const int wu = __builtin_amdgcn_readfirstlane(w);
        const int2 wk = work[wu];
        const BatchCase& bc = cases[wk.x];
        const int in = stage == 0 ? 0 : (stage == 1 ? 1 : 2);
        const int out = stage == 0 ? 1 : (stage == 1 ? 2 : 0);
        const float* U[4] = {(const float*)bc.P[in][0], (const float*)bc.P[in][1], (const float*)bc.P[in][2], (const float*)bc.P[in][3]};
        const float* U0[4] = {(const float*)bc.P[0][0], (const float*)bc.P[0][1], (const float*)bc.P[0][2], (const float*)bc.P[0][3]};
        float* O[4] = {(float*)bc.P[out][0], (float*)bc.P[out][1], (float*)bc.P[out][2], (float*)bc.P[out][3]};
        stage_f32_strip<FLUX, LIM>(bc.a, wk.y, U, U0, O, stage, parity);
    }
}

hipError_t launch_stage_f32_batch(const BatchCase* cases, const int2* work, int n_work, int flux, int lim, int stage,
                                  int parity, int blocks, int waves, hipStream_t st) {
    dim3 grid(blocks), block(WAVE * waves);
#define EFVM_F32B_CASE(F)                                                                                        \
    if (lim == LIM_NONE) stage_f32_batch_kernel<F, LIM_NONE><<<grid, block, 0, st>>>(cases, work, n_work, stage, parity);      \
    else if (lim == LIM_MINMOD) stage_f32_batch_kernel<F, LIM_MINMOD><<<grid, block, 0, st>>>(cases, work, n_work, stage, parity); \
    else stage_f32_batch_kernel<F, LIM_MC><<<grid, block, 0, st>>>(cases, work, n_work, stage, parity);
    switch (flux) {
    case FLUX_RUSANOV: EFVM_F32B_CASE(FLUX_RUSANOV); break;
    case FLUX_HLLC: EFVM_F32B_CASE(FLUX_HLLC); break;
    default: EFVM_F32B_CASE(FLUX_HLLC_TEXTBOOK); break;
    }
#undef EFVM_F32B_CASE
    return hipGetLastError();
}

hipError_t launch_stage_f32(const StageArgs& a, int block, hipStream_t st) {
    int waves = block / WAVE;
    int nstrips = a.n_strips_x * a.n_strips_y;
    dim3 grid((nstrips + waves - 1) / waves);
    const int lim = a.muscl ? a.limiter : LIM_NONE;
#define EFVM_F32_CASE(F)                                                                   \
    if (lim == LIM_NONE) stage_f32_kernel<F, LIM_NONE><<<grid, block, 0, st>>>(a);          \
    else if (lim == LIM_MINMOD) stage_f32_kernel<F, LIM_MINMOD><<<grid, block, 0, st>>>(a); \
    else stage_f32_kernel<F, LIM_MC><<<grid, block, 0, st>>>(a);
    switch (a.flux) {
    case FLUX_RUSANOV: EFVM_F32_CASE(FLUX_RUSANOV); break;
    case FLUX_HLLC: EFVM_F32_CASE(FLUX_HLLC); break;
    default: EFVM_F32_CASE(FLUX_HLLC_TEXTBOOK); break;
    }
#undef EFVM_F32_CASE
    return hipGetLastError();
}

}  // namespace efvm
