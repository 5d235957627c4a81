// efvm_stage_f64.hip -- FP64 parity path (compiled with -ffp-contract=off).
//
// The stage kernel reproduces the reference's per-cell arithmetic op for op:
// face states, Riemann flux and the two-sided scatter in the order the
// structured face sweep visits faces (crates/cfd_solver/src/lib.rs:410-446):
//   rhs = 0 (+G_s if j>0) (+F_w if i>0) -F_e (+F_w if i==0) -G_n (+G_s if j==0)
//         -F_zhi +F_zlo            [each term F*(area/volume)]
// including the two z faces of the one-cell-thick 2-D mesh, whose slip-wall
// ghosts contribute rounding-level terms (crates/cfd_boundary/src/lib.rs:246-257).
// Then viscous (cfd_solver:463-560), body force, solid-rhs zero (cfd_ibm:293-301)
// and the RK substep (integrate.rs:75-101).
#include "efvm_stage.h"

namespace efvm {

using D5 = C5<double>;

__device__ __forceinline__ D5 shfl_up5(const D5& v) { return {shfl_up1(v.r), shfl_up1(v.ru), shfl_up1(v.rv), v.rw, shfl_up1(v.re)}; }
__device__ __forceinline__ D5 shfl_dn5(const D5& v) { return {shfl_dn1(v.r), shfl_dn1(v.ru), shfl_dn1(v.rv), v.rw, shfl_dn1(v.re)}; }
__device__ __forceinline__ D5 as5(const Cell<double>& c) { return {c.r, c.ru, c.rv, 0.0, c.re}; }
__device__ __forceinline__ D5 lim5(int lim, const D5& f, const D5& b) {
    return {ref::limit(lim, f.r, b.r), ref::limit(lim, f.ru, b.ru), ref::limit(lim, f.rv, b.rv),
            ref::limit(lim, f.rw, b.rw), ref::limit(lim, f.re, b.re)};
}
__device__ __forceinline__ D5 zero5() { return {0.0, 0.0, 0.0, 0.0, 0.0}; }

struct PrimTV { double u, v, t; };
__device__ __forceinline__ PrimTV prim_tv(const StageArgs& a, const D5& c) {
    P6<double> p = ref::cons_to_prim(c, a.gamma, a.r_gas);
    return {p.u, p.v, p.t};
}

template <int FLUX, int MUSCL>
__global__ __launch_bounds__(256) void stage_f64_kernel(StageArgs a) {
    const int lane = lane_id();
    const int s = strip_index(a, blockDim.x / WAVE);
    if (s >= a.n_strips_x * a.n_strips_y) return;
    const int nx = a.nx, ny = a.ny;
    const int sxi = s % a.n_strips_x, syi = s / a.n_strips_x;
    const int c = sxi * STRIP_OUT - 2 + lane;
    const int j0 = syi * a.strip_h;
    const int j1 = min(j0 + a.strip_h, ny);
    const double* U[4] = {(const double*)a.uin[0], (const double*)a.uin[1], (const double*)a.uin[2], (const double*)a.uin[3]};
    const double* U0[4] = {(const double*)a.u0[0], (const double*)a.u0[1], (const double*)a.u0[2], (const double*)a.u0[3]};
    double* O[4] = {(double*)a.uout[0], (double*)a.uout[1], (double*)a.uout[2], (double*)a.uout[3]};
    const double g = a.gamma;
    const int lim = a.limiter;
    const double dt = stage_dt(a);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (a.stage == 0) { a.clock->dt = dt; a.clock->lam[a.parity ^ 1] = 0.0; }
    }

    // per-column reconstruction rules (see header comment of efvm_stage.h)
    const bool col_ghost = (c < 0 || c >= nx);
    const bool seam_rec = a.seam_mode == 1;
    const bool xW_rec = MUSCL && (col_ghost ? (a.per_x && seam_rec) : !(c == 0 && !a.per_x));
    const bool xE_rec = MUSCL && (col_ghost ? (a.per_x && seam_rec) : !(c == nx - 1 && !a.per_x));
    const bool out_lane = lane >= 2 && lane < WAVE - 2 && c < nx;

    D5 um = as5(load_cell(a, U, c, j0 - 2));
    D5 uc = as5(load_cell(a, U, c, j0 - 1));
    D5 up = as5(load_cell(a, U, c, j0));
    // viscous rolling window: primitives and solid flags of rows jc-1, jc, jc+1
    PrimTV pm{}, pc{}, pp{};
    int sm = 0, sc = 0, sp = 0;
    const bool visc = a.viscous != 0;
    const bool ibm = a.solid != nullptr;
    auto solid_at = [&](int cc, int rr) -> int {
        if (!ibm || cc < 0 || cc >= nx || rr < 0 || rr >= ny) return 0;
        return a.solid[(long)rr * nx + cc];
    };
    if (visc) { pm = prim_tv(a, um); pc = prim_tv(a, uc); pp = prim_tv(a, up); }
    if (ibm) { sm = solid_at(c, j0 - 2); sc = solid_at(c, j0 - 1); sp = solid_at(c, j0); }

    D5 hi_prev = zero5(), G_prev = zero5(), Fw_prev = zero5(), Fe_prev = zero5();
    double vis_prev[4] = {0, 0, 0, 0};
    double lam_max = 0.0;
    long long bad = 0;

    for (int jc = j0 - 1; jc <= j1; ++jc) {
        // ---- y states of row jc and the y face jc-1/2 ----
        const bool row_ghost = (jc < 0 || jc >= ny);
        const bool yS_rec = MUSCL && (row_ghost ? (a.per_y && seam_rec) : !(jc == 0 && !a.per_y));
        const bool yN_rec = MUSCL && (row_ghost ? (a.per_y && seam_rec) : !(jc == ny - 1 && !a.per_y));
        D5 lo_st = uc, hi_st = uc;
        if (MUSCL) {
            D5 sy = lim5(lim, ref::sub(up, uc), ref::sub(uc, um));
            if (yS_rec) lo_st = ref::add(uc, ref::mul(sy, -0.5));
            if (yN_rec) hi_st = ref::add(uc, ref::mul(sy, 0.5));
        }
        D5 G = zero5();
        if (jc >= j0) G = ref::num_flux(FLUX, hi_prev, lo_st, g, 1);

        // ---- x faces and viscous terms of row jc ----
        D5 Fw = zero5(), Fe = zero5();
        double vis[4] = {0, 0, 0, 0};
        if (jc >= j0 && jc < j1) {
            D5 ul = shfl_up5(uc), ur = shfl_dn5(uc);
            D5 W_st = uc, E_st = uc;
            if (MUSCL) {
                D5 sxs = lim5(lim, ref::sub(ur, uc), ref::sub(uc, ul));
                if (xW_rec) W_st = ref::add(uc, ref::mul(sxs, -0.5));
                if (xE_rec) E_st = ref::add(uc, ref::mul(sxs, 0.5));
            }
            D5 El = shfl_up5(E_st);
            Fw = ref::num_flux(FLUX, El, W_st, g, 0);
            Fe = shfl_dn5(Fw);
            if (visc) {
                // neighbour values with the no-slip/adiabatic substitution at solids
                PrimTV pl_{shfl_up1(pc.u), shfl_up1(pc.v), shfl_up1(pc.t)};
                PrimTV pr_{shfl_dn1(pc.u), shfl_dn1(pc.v), shfl_dn1(pc.t)};
                int sl_ = shfl_up1(sc), sr_ = shfl_dn1(sc);
                if (!sc && !col_ghost) {
                    double mu = a.visc_const ? a.mu : ref::sutherland(pc.t);
                    double kap = ref::conductivity(mu, a.gamma, a.r_gas);
                    double q[4][3];   // [xp, xm, yp, ym] x [u, v, t]
                    const PrimTV* nb[4] = {&pr_, &pl_, &pp, &pm};
                    const int nbs[4] = {sr_, sl_, sp, sm};
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        if (nbs[t]) { q[t][0] = 0.0; q[t][1] = 0.0; q[t][2] = pc.t; }
                        else { q[t][0] = nb[t]->u; q[t][1] = nb[t]->v; q[t][2] = nb[t]->t; }
                    }
                    const double yim = a.yw[jc], yip = a.yw[ny + jc], yis = a.yw[2 * ny + jc];
                    double vc[3] = {pc.u, pc.v, pc.t};
                    double lap[3];
#pragma unroll
                    for (int m = 0; m < 3; m++) {
                        double d2x = (q[0][m] - 2.0 * vc[m] + q[1][m]) * a.dxi2;
                        double d2y = ((q[2][m] - vc[m]) * yip - (vc[m] - q[3][m]) * yim) * yis;
                        lap[m] = d2x + d2y;
                    }
                    double fx = mu * lap[0], fy = mu * lap[1];
                    double en = vc[0] * fx + vc[1] * fy + kap * lap[2];
                    vis[0] = fx; vis[1] = fy; vis[2] = en; vis[3] = 1.0;
                }
            }
        }

        // ---- output row jc-1 ----
        if (jc - 1 >= j0 && out_lane) {
            const int j = jc - 1;
            const int i = c;
            const D5 u = um;   // U(j)
            D5 rhs = zero5();
            D5 tx, ty;
            if (j > 0) rhs = ref::add(rhs, ref::mul(G_prev, a.inv_y));
            if (i > 0) rhs = ref::add(rhs, ref::mul(Fw_prev, a.inv_x));
            rhs = ref::sub(rhs, ref::mul(Fe_prev, a.inv_x));
            if (i == 0) rhs = ref::add(rhs, ref::mul(Fw_prev, a.inv_x));
            rhs = ref::sub(rhs, ref::mul(G, a.inv_y));
            if (j == 0) rhs = ref::add(rhs, ref::mul(G_prev, a.inv_y));
            {   // z faces: slip-wall mirror of the cell itself (rho_w -> -rho_w)
                D5 mir = u; mir.rw = -u.rw;
                D5 fzh = ref::num_flux(FLUX, u, mir, g, 2);
                D5 fzl = ref::num_flux(FLUX, mir, u, g, 2);
                rhs = ref::sub(rhs, ref::mul(fzh, a.inv_z));
                rhs = ref::add(rhs, ref::mul(fzl, a.inv_z));
            }
            (void)tx; (void)ty;
            if (vis_prev[3] != 0.0) {
                rhs.r = rhs.r + 0.0;
                rhs.ru = rhs.ru + vis_prev[0];
                rhs.rv = rhs.rv + vis_prev[1];
                rhs.re = rhs.re + vis_prev[2];
            }
            if (a.body_force) {
                rhs.ru += u.r * a.gx;
                rhs.re += a.gx * u.ru;
            }
            const long q = (long)j * nx + i;
            if (ibm && a.solid[q]) rhs = zero5();
            D5 o;
            D5 step = ref::add(u, ref::mul(rhs, dt));
            if (a.euler || a.stage == 0) o = step;
            else {
                D5 u0 = {U0[0][q], U0[1][q], U0[2][q], 0.0, U0[3][q]};
                if (a.stage == 1) o = ref::add(ref::mul(u0, 0.75), ref::mul(step, 0.25));
                else o = ref::add(ref::mul(u0, 1.0 / 3.0), ref::mul(step, 2.0 / 3.0));
            }
            O[0][q] = o.r; O[1][q] = o.ru; O[2][q] = o.rv; O[3][q] = o.re;
            if (a.fuse_cfl) {
                if (o.r > 0.0) {
                    double l = cfl_lambda(a, o.r, o.ru, o.rv, o.re);
                    if (l == l) lam_max = fmax(lam_max, l);
                    else bad++;
                } else bad++;
            }
        }

        // ---- roll the window ----
        hi_prev = hi_st; G_prev = G; Fw_prev = Fw; Fe_prev = Fe;
#pragma unroll
        for (int m = 0; m < 4; m++) vis_prev[m] = vis[m];
        um = uc; uc = up;
        if (visc) { pm = pc; pc = pp; }
        if (ibm) { sm = sc; sc = sp; }
        if (jc + 2 <= j1 + 1) {
            up = as5(load_cell(a, U, c, jc + 2));
            if (visc) pp = prim_tv(a, up);
            if (ibm) sp = solid_at(c, jc + 2);
        }
    }
    if (a.fuse_cfl) {
        double w = wave_max(lam_max);
        long long b = wave_sum_ll(bad);
        if (lane == 0) {
            if (w > 0.0) atomic_max_pos_double(&a.clock->lam[a.parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (a.euler || a.stage == 2)) {
        a.clock->time += dt;
        a.clock->step += 1;
    }
}

// CFL scan over the interior (crates/cfd_solver/src/lib.rs:626-677) into lam[target].
template <class T>
__global__ __launch_bounds__(256) void cfl_kernel(StageArgs a, const T* __restrict__ r, const T* __restrict__ ru,
                                                  const T* __restrict__ rv, const T* __restrict__ re, double* target) {
    long n = (long)a.nx * a.ny;
    double m = 0.0;
    long long bad = 0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double rr = r[q];
        if (rr <= 0.0 || !(rr == rr)) { bad++; continue; }
        double l = cfl_lambda(a, rr, (double)ru[q], (double)rv[q], (double)re[q]);
        if (l == l) m = fmax(m, l); else bad++;
    }
    m = wave_max(m);
    bad = wave_sum_ll(bad);
    if (lane_id() == 0) {
        if (m > 0.0) atomic_max_pos_double(target, m);
        if (bad) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)bad);
    }
}

__global__ void reset_double(double* target) { *target = 0.0; }
__global__ void clock_set(DevClock* clk, double dt, double time, long long step) {
    clk->dt = dt; clk->time = time; clk->step = step; clk->bad = 0; clk->lam[0] = 0.0; clk->lam[1] = 0.0;
}

// ---- immersed boundary: image-point trilinear mirror (cfd_ibm:106-149, 258-289) ----
struct IbmArgs {
    int nx, ny;
    double lo_x, lo_y, dx, dy, dz;
    double gamma, r_gas;
    int per_x, per_y;
    const uint8_t* solid;
};

// locate(): s = (p - lo)/h - 0.5, i0 = floor(s), t = s - i0; z (nz = 1) handled by the caller.
template <class T>
__device__ bool sample_trilinear(const StageArgs& sa, const IbmArgs& a, const T* const* U, const double* pt, P6<double>& out) {
    double sxv = (pt[0] - a.lo_x) / a.dx - 0.5, syv = (pt[1] - a.lo_y) / a.dy - 0.5, szv = (pt[2] - 0.0) / a.dz - 0.5;
    double fx = floor(sxv), fy = floor(syv), fz = floor(szv);
    int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    if (i0 < -1 || i0 > a.nx - 1 || j0 < -1 || j0 > a.ny - 1 || k0 < -1 || k0 > 0) return false;
    double t[3] = {sxv - fx, syv - fy, szv - fz};
    double acc[6] = {0, 0, 0, 0, 0, 0}, tw = 0.0;
    for (int di = 0; di < 2; di++) {
        double wx = di == 0 ? 1.0 - t[0] : t[0];
        for (int dj = 0; dj < 2; dj++) {
            double wy = dj == 0 ? 1.0 - t[1] : t[1];
            for (int dk = 0; dk < 2; dk++) {
                double wz = dk == 0 ? 1.0 - t[2] : t[2];
                int ci = i0 + di, cj = j0 + dj, ck = k0 + dk;   // ck: 0 = the cell layer, else a z ghost
                bool interior = ci >= 0 && ci < a.nx && cj >= 0 && cj < a.ny;
                if (interior && a.solid && a.solid[(long)cj * a.nx + ci]) continue;
                Cell<T> cl = load_cell(sa, U, ci, cj);
                C5<double> u = {double(cl.r), double(cl.ru), double(cl.rv), 0.0, double(cl.re)};
                if (ck != 0) u.rw = -u.rw;   // z slip ghost
                if (!(u.r == u.r) || isinf(u.r) || u.r <= 0.0) continue;
                double w = wx * wy * wz;
                P6<double> p = ref::cons_to_prim(u, a.gamma, a.r_gas);
                acc[0] += w * p.r; acc[1] += w * p.u; acc[2] += w * p.v; acc[3] += w * p.w; acc[4] += w * p.p; acc[5] += w * p.t;
                tw += w;
            }
        }
    }
    if (tw < 1e-12) return false;
    double inv = 1.0 / tw;
    out = {acc[0] * inv, acc[1] * inv, acc[2] * inv, acc[3] * inv, acc[4] * inv, acc[5] * inv};
    return true;
}

// Mirror: u_ghost = prim_to_cons(rho_img, 2*0 - u_img, p_img).  Sampling skips solid
// cells and writes only solid cells, so ghosts are independent of one another.  `pre`
// (4 planes of n) keeps each ghost's value from before the mirror: the boundary fill runs
// before the mirror, so ghost cells filled from a solid cell hold that value.
template <class T>
__global__ void ibm_mirror_kernel(StageArgs sa, IbmArgs a, int n, const int* __restrict__ idx, const double* __restrict__ img,
                                  T* __restrict__ pre) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const T* U[4] = {(const T*)sa.uin[0], (const T*)sa.uin[1], (const T*)sa.uin[2], (const T*)sa.uin[3]};
    if (pre) {
        const long k0 = idx[q];
        for (int m = 0; m < 4; m++) pre[(long)m * n + q] = U[m][k0];
    }
    P6<double> im;
    if (!sample_trilinear<T>(sa, a, U, img + 3 * q, im)) return;
    P6<double> m = {im.r, 2.0 * 0.0 - im.u, 2.0 * 0.0 - im.v, 2.0 * 0.0 - im.w, im.p, im.t};
    C5<double> c = ref::prim_to_cons(m, a.gamma);
    T* W[4] = {(T*)sa.uin[0], (T*)sa.uin[1], (T*)sa.uin[2], (T*)sa.uin[3]};
    long k = idx[q];
    W[0][k] = T(c.r); W[1][k] = T(c.ru); W[2][k] = T(c.rv); W[3][k] = T(c.re);
}

// Marker traction integral (cfd_ibm:442-480), one wave, serial sum in marker order
// so the FP64 result does not depend on scheduling.
template <class T>
__global__ void marker_force_kernel(StageArgs sa, IbmArgs a, int n, const double* __restrict__ mk, double off,
                                    int visc_const, double mu_c, double* out) {
    __shared__ double part[64][6];
    int l = threadIdx.x;
    const T* U[4] = {(const T*)sa.uin[0], (const T*)sa.uin[1], (const T*)sa.uin[2], (const T*)sa.uin[3]};
    // each lane evaluates its markers' contributions; lane 0 then sums in marker order
    for (int base = 0; base < n; base += 64) {
        int q = base + l;
        double c6[6] = {0, 0, 0, 0, 0, 0};
        bool ok = false;
        if (q < n) {
            const double* m = mk + 7 * q;
            double sp[3] = {m[0] + off * m[3], m[1] + off * m[4], m[2] + off * m[5]};
            P6<double> im;
            if (sample_trilinear<T>(sa, a, U, sp, im) && isfinite(im.p) && isfinite(im.t) && isfinite(im.u) &&
                isfinite(im.v) && isfinite(im.w)) {
                double us[3] = {im.u - 0.0, im.v - 0.0, im.w - 0.0};
                double udn = us[0] * m[3] + us[1] * m[4] + us[2] * m[5];
                double mu = visc_const ? mu_c : ref::sutherland(im.t);
                for (int d = 0; d < 3; d++) {
                    double ut = us[d] - udn * m[3 + d];
                    c6[d] = -im.p * m[3 + d] * m[6];
                    c6[3 + d] = mu * ut / off * m[6];
                }
                ok = true;
            }
        }
        for (int d = 0; d < 6; d++) part[l][d] = ok ? c6[d] : 0.0;
        __shared__ int okf[64];
        okf[l] = ok;
        __syncthreads();
        if (l == 0) {
            for (int k = 0; k < 64 && base + k < n; k++)
                if (okf[k]) for (int d = 0; d < 6; d++) out[d] += part[k][d];
        }
        __syncthreads();
    }
}
__global__ void zero6(double* out) { if (threadIdx.x < 6) out[threadIdx.x] = 0.0; }

// Staircase wall force (cfd_ibm:309-362): pressure -p n A and wall shear mu u_t / (dn/2) A
// summed over every solid-fluid face.  The ghost list holds every solid cell in row-major
// order (on the one-cell-thick mesh each has fluid z neighbours), so lane l evaluates the six
// faces of ghost base + l and lane 0 adds the terms in cell order and face order x+, x-, y+,
// y-, z+, z-: the FP64 sum is the serial one.
//
// Neighbours inside the grid come from the state U.  A ghost cell holds what the substages
// since the last step boundary made of it: each substage fills it from its input (a z ghost is
// the slip image of the cell, an x/y ghost the periodic image or the boundary condition of its
// owner cell, taken before the IBM mirror overwrites solid cells), and each update then combines
// it like an interior cell with a zero right-hand side.  So the ghost is the fill of src[0]
// (one substage since the step boundary), or w0 fill(src[0]) + w1 fill(src[1]) (RK3: 3/4, 1/4
// after two substages, 1/3, 2/3 after the step); with no substage run it is taken from U.
struct WallSources {
    int n_src;                 // 0, 1 or 2
    const void* set[2][4];     // the substage inputs whose fills the ghosts combine
    const void* pre[2];        // their solid cells before the mirror (4 planes of the ghost count)
    double w[2];
};

template <class T>
__device__ Cell<T> pre_mirror(const StageArgs& sa, const T* const* X, const T* pre, const int* idx, int n, int col, int row) {
    const long k = (long)row * sa.nx + col;
    if (sa.solid && sa.solid[k]) {
        int lo = 0, hi = n - 1;   // idx ascends: binary search for the cell's ghost slot
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (idx[mid] < k) lo = mid + 1; else hi = mid;
        }
        return {pre[lo], pre[(long)n + lo], pre[2L * n + lo], pre[3L * n + lo]};
    }
    return load_plain(sa, X, col, row);
}

// The fill from one source of the ghost across face f of cell (i, j) (ghost slot q).
template <class T>
__device__ C5<double> ghost_fill(const StageArgs& sa, const T* const* X, const T* pre, const int* idx, int n, int q, int i,
                                 int j, int f) {
    Cell<T> cl;
    if ((f >> 1) == 2) {
        // slip image of the cell itself: only rho_w changes sign, and it is zero on this mesh
        cl = {pre[q], pre[(long)n + q], pre[2L * n + q], pre[3L * n + q]};
    } else {
        const int ci = i + (f == 0 ? 1 : f == 1 ? -1 : 0), cj = j + (f == 2 ? 1 : f == 3 ? -1 : 0);
        const GhostOwner g = ghost_owner(sa, ci, cj);
        const Cell<T> own = pre_mirror<T>(sa, X, pre, idx, n, g.col, g.row);
        cl = g.side < 0 ? own : ghost_of_owner(sa, g.side, own, g.row, g.col);
    }
    return {double(cl.r), double(cl.ru), double(cl.rv), 0.0, double(cl.re)};
}

template <class T>
__global__ void wall_force_kernel(StageArgs sa, IbmArgs a, int n, const int* __restrict__ idx, WallSources ws,
                                  int visc_const, double mu_c, double* out) {
    __shared__ double term[64][6][3];   // per face: pressure term, then the two tangential shear terms
    __shared__ unsigned char fluid[64][6];
    const int l = threadIdx.x;
    const T* U[4] = {(const T*)sa.uin[0], (const T*)sa.uin[1], (const T*)sa.uin[2], (const T*)sa.uin[3]};
    const T* X0[4] = {(const T*)ws.set[0][0], (const T*)ws.set[0][1], (const T*)ws.set[0][2], (const T*)ws.set[0][3]};
    const T* X1[4] = {(const T*)ws.set[1][0], (const T*)ws.set[1][1], (const T*)ws.set[1][2], (const T*)ws.set[1][3]};
    const double fa[3] = {a.dy * a.dz, a.dx * a.dz, a.dx * a.dy}, dn[3] = {a.dx, a.dy, a.dz};
    double fp[3] = {0.0, 0.0, 0.0}, fv[3] = {0.0, 0.0, 0.0};
    for (int base = 0; base < n; base += 64) {
        const int q = base + l;
        if (q < n) {
            const long k = idx[q];
            const int i = (int)(k % a.nx), j = (int)(k / a.nx);
            for (int f = 0; f < 6; f++) {
                const int dim = f >> 1;
                const double sg = (f & 1) ? -1.0 : 1.0;
                const int ci = i + (f == 0 ? 1 : f == 1 ? -1 : 0), cj = j + (f == 2 ? 1 : f == 3 ? -1 : 0);
                const bool inside = dim < 2 && ci >= 0 && ci < a.nx && cj >= 0 && cj < a.ny;
                C5<double> u;
                bool fl = true;
                if (inside) {
                    fl = !a.solid[(long)cj * a.nx + ci];
                    const Cell<T> cl = load_plain(sa, U, ci, cj);
                    u = {double(cl.r), double(cl.ru), double(cl.rv), 0.0, double(cl.re)};
                } else if (ws.n_src == 0) {
                    const Cell<T> cl = dim == 2 ? load_plain(sa, U, i, j) : load_cell(sa, U, ci, cj);
                    u = {double(cl.r), double(cl.ru), double(cl.rv), 0.0, double(cl.re)};
                } else {
                    u = ghost_fill<T>(sa, X0, (const T*)ws.pre[0], idx, n, q, i, j, f);
                    if (ws.n_src == 2) {
                        // the update of a ghost: w0 u0 + w1 (u + 0 dt), component by component
                        const C5<double> v = ghost_fill<T>(sa, X1, (const T*)ws.pre[1], idx, n, q, i, j, f);
                        u = {u.r * ws.w[0] + v.r * ws.w[1], u.ru * ws.w[0] + v.ru * ws.w[1], u.rv * ws.w[0] + v.rv * ws.w[1],
                             0.0, u.re * ws.w[0] + v.re * ws.w[1]};
                    }
                }
                fluid[l][f] = fl;
                if (!fl) continue;
                term[l][f][0] = ref::pressure(u, a.gamma) * sg * fa[dim];
                const P6<double> pr = ref::cons_to_prim(u, a.gamma, a.r_gas);
                const double mu = visc_const ? mu_c : ref::sutherland(pr.t);
                const double vel[3] = {pr.u, pr.v, pr.w};
                for (int td = 0, t = 1; td < 3; td++)
                    if (td != dim) term[l][f][t++] = mu * vel[td] / (0.5 * dn[dim]) * fa[dim];
            }
        }
        __syncthreads();
        if (l == 0) {
            for (int c = 0; c < 64 && base + c < n; c++)
                for (int f = 0; f < 6; f++) {
                    if (!fluid[c][f]) continue;
                    const int dim = f >> 1;
                    fp[dim] -= term[c][f][0];
                    for (int td = 0, t = 1; td < 3; td++)
                        if (td != dim) fv[td] += term[c][f][t++];
                }
        }
        __syncthreads();
    }
    if (l == 0)
        for (int d = 0; d < 3; d++) { out[d] = fp[d]; out[3 + d] = fv[d]; }
}

// Outlet sponge (examples/cylinder_drag_validation/main.rs:186-229), after a completed step.
template <class T>
__global__ void sponge_kernel(StageArgs a, double x0, double x1, double sig, double um, double h, double rho, double p) {
    long n = (long)a.nx * a.ny;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int i = (int)(q % a.nx), j = (int)(q / a.nx);
    double cx = a.lo_x + ((double)i + 0.5) * a.dx, cy = a.lo_y + ((double)j + 0.5) * a.dy;
    if (cx < x0) return;
    double dt = a.clock->dt;
    double inv_w = 1.0 / (x1 - x0);
    double sc = (cx - x0) * inv_w;
    sc = sc < 0.0 ? 0.0 : (sc > 1.0 ? 1.0 : sc);
    double f = fmin(sig * sc * sc * dt, 1.0);
    double uy = 4.0 * um * cy * (h - cy) / (h * h);
    P6<double> pv = {rho, uy, 0.0, 0.0, p, p / (rho * 287.058)};
    C5<double> tg = ref::prim_to_cons(pv, a.gamma);
    T* W[4] = {(T*)a.uout[0], (T*)a.uout[1], (T*)a.uout[2], (T*)a.uout[3]};
    double r = W[0][q], ru = W[1][q], rv = W[2][q], re = W[3][q];
    r += f * (tg.r - r); ru += f * (tg.ru - ru); rv += f * (tg.rv - rv); re += f * (tg.re - re);
    W[0][q] = T(r); W[1][q] = T(ru); W[2][q] = T(rv); W[3][q] = T(re);
}

// FP64 cell-volume-weighted sums of the 4 stored components: per-block partials
// (fixed order) then a single-block fixed-order combine -> deterministic.
template <class T>
__global__ __launch_bounds__(256) void sums_partial_kernel(StageArgs a, double* part) {
    __shared__ double sh[256][4];
    long n = (long)a.nx * a.ny;
    const T* U[4] = {(const T*)a.uin[0], (const T*)a.uin[1], (const T*)a.uin[2], (const T*)a.uin[3]};
    double s4[4] = {0, 0, 0, 0};
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        for (int m = 0; m < 4; m++) s4[m] += (double)U[m][q];
    for (int m = 0; m < 4; m++) sh[threadIdx.x][m] = s4[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) for (int m = 0; m < 4; m++) sh[threadIdx.x][m] += sh[threadIdx.x + w][m];
        __syncthreads();
    }
    if (threadIdx.x == 0) for (int m = 0; m < 4; m++) part[4 * blockIdx.x + m] = sh[0][m];
}
__global__ void sums_final_kernel(const double* part, int nb, double vol, double* out) {
    if (threadIdx.x != 0) return;
    double s[4] = {0, 0, 0, 0};
    for (int b = 0; b < nb; b++) for (int m = 0; m < 4; m++) s[m] += part[4 * b + m];
    for (int m = 0; m < 4; m++) out[m] = s[m] * vol;
}

}  // namespace efvm

// ---------------------------------------------------------------------------
// host launchers (C++ linkage, used by efvm_api.cpp)
// ---------------------------------------------------------------------------
namespace efvm {

hipError_t launch_stage_f64(const StageArgs& a, int block, hipStream_t st) {
    int waves = block / WAVE;
    int nstrips = a.n_strips_x * a.n_strips_y;
    dim3 grid((nstrips + waves - 1) / waves);
    if (a.muscl) {
        switch (a.flux) {
        case FLUX_RUSANOV: stage_f64_kernel<FLUX_RUSANOV, 1><<<grid, block, 0, st>>>(a); break;
        case FLUX_HLLC: stage_f64_kernel<FLUX_HLLC, 1><<<grid, block, 0, st>>>(a); break;
        default: stage_f64_kernel<FLUX_HLLC_TEXTBOOK, 1><<<grid, block, 0, st>>>(a); break;
        }
    } else {
        switch (a.flux) {
        case FLUX_RUSANOV: stage_f64_kernel<FLUX_RUSANOV, 0><<<grid, block, 0, st>>>(a); break;
        case FLUX_HLLC: stage_f64_kernel<FLUX_HLLC, 0><<<grid, block, 0, st>>>(a); break;
        default: stage_f64_kernel<FLUX_HLLC_TEXTBOOK, 0><<<grid, block, 0, st>>>(a); break;
        }
    }
    return hipGetLastError();
}

template <class T>
hipError_t launch_cfl(const StageArgs& a, const void* const* planes, double* target, hipStream_t st) {
    long n = (long)a.nx * a.ny;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    reset_double<<<1, 1, 0, st>>>(target);
    cfl_kernel<T><<<blocks, 256, 0, st>>>(a, (const T*)planes[0], (const T*)planes[1], (const T*)planes[2], (const T*)planes[3], target);
    return hipGetLastError();
}
template hipError_t launch_cfl<double>(const StageArgs&, const void* const*, double*, hipStream_t);
template hipError_t launch_cfl<float>(const StageArgs&, const void* const*, double*, hipStream_t);

hipError_t launch_clock_set(DevClock* clk, double dt, double time, long long step, hipStream_t st) {
    clock_set<<<1, 1, 0, st>>>(clk, dt, time, step);
    return hipGetLastError();
}

template <class T>
hipError_t launch_ibm_mirror(const StageArgs& sa, const IbmArgs& a, int n, const int* idx, const double* img, void* pre, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    ibm_mirror_kernel<T><<<(n + 127) / 128, 128, 0, st>>>(sa, a, n, idx, img, (T*)pre);
    return hipGetLastError();
}
template hipError_t launch_ibm_mirror<double>(const StageArgs&, const IbmArgs&, int, const int*, const double*, void*, hipStream_t);
template hipError_t launch_ibm_mirror<float>(const StageArgs&, const IbmArgs&, int, const int*, const double*, void*, hipStream_t);

template <class T>
hipError_t launch_wall_force(const StageArgs& sa, const IbmArgs& a, int n, const int* idx, int n_src, void* const* set0,
                             void* const* set1, const void* pre0, const void* pre1, double w0, double w1, int visc_const,
                             double mu, double* out, hipStream_t st) {
    WallSources ws;
    ws.n_src = n_src;
    for (int m = 0; m < 4; m++) { ws.set[0][m] = set0 ? set0[m] : nullptr; ws.set[1][m] = set1 ? set1[m] : nullptr; }
    ws.pre[0] = pre0; ws.pre[1] = pre1;
    ws.w[0] = w0; ws.w[1] = w1;
    wall_force_kernel<T><<<1, 64, 0, st>>>(sa, a, n, idx, ws, visc_const, mu, out);
    return hipGetLastError();
}
template hipError_t launch_wall_force<double>(const StageArgs&, const IbmArgs&, int, const int*, int, void* const*, void* const*,
                                              const void*, const void*, double, double, int, double, double*, hipStream_t);
template hipError_t launch_wall_force<float>(const StageArgs&, const IbmArgs&, int, const int*, int, void* const*, void* const*,
                                             const void*, const void*, double, double, int, double, double*, hipStream_t);

template <class T>
hipError_t launch_marker_force(const StageArgs& sa, const IbmArgs& a, int n, const double* mk, double off, int visc_const,
                               double mu, double* out, hipStream_t st) {
    zero6<<<1, 64, 0, st>>>(out);
    marker_force_kernel<T><<<1, 64, 0, st>>>(sa, a, n, mk, off, visc_const, mu, out);
    return hipGetLastError();
}
template hipError_t launch_marker_force<double>(const StageArgs&, const IbmArgs&, int, const double*, double, int, double, double*, hipStream_t);
template hipError_t launch_marker_force<float>(const StageArgs&, const IbmArgs&, int, const double*, double, int, double, double*, hipStream_t);

template <class T>
hipError_t launch_sponge(const StageArgs& a, double x0, double x1, double sig, double um, double h, double rho, double p, hipStream_t st) {
    long n = (long)a.nx * a.ny;
    sponge_kernel<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a, x0, x1, sig, um, h, rho, p);
    return hipGetLastError();
}
template hipError_t launch_sponge<double>(const StageArgs&, double, double, double, double, double, double, double, hipStream_t);
template hipError_t launch_sponge<float>(const StageArgs&, double, double, double, double, double, double, double, hipStream_t);

template <class T>
hipError_t launch_sums(const StageArgs& a, double* part, int nb, double vol, double* out, hipStream_t st) {
    sums_partial_kernel<T><<<nb, 256, 0, st>>>(a, part);
    sums_final_kernel<<<1, 64, 0, st>>>(part, nb, vol, out);
    return hipGetLastError();
}
template hipError_t launch_sums<double>(const StageArgs&, double*, int, double, double*, hipStream_t);
template hipError_t launch_sums<float>(const StageArgs&, double*, int, double, double*, hipStream_t);

// The per-block partials alone (4 doubles per block into `part`), for callers that finish the sum themselves.
template <class T>
hipError_t launch_sums_partial(const StageArgs& a, double* part, int nb, hipStream_t st) {
    sums_partial_kernel<T><<<nb, 256, 0, st>>>(a, part);
    return hipGetLastError();
}
template hipError_t launch_sums_partial<double>(const StageArgs&, double*, int, hipStream_t);
template hipError_t launch_sums_partial<float>(const StageArgs&, double*, int, hipStream_t);

}  // namespace efvm

// ---- cylinder mask refresh: newly solid cells take the stagnant state (cfd_ibm:155-183) ----
namespace efvm {
template <class T>
__global__ void ibm_flip_kernel(T* r, T* ru, T* rv, T* re, const uint8_t* __restrict__ solid, long n, const double* __restrict__ st) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || !solid[q]) return;
    r[q] = T(st[0]); ru[q] = T(st[1]); rv[q] = T(st[2]); re[q] = T(st[4]);
}
hipError_t launch_ibm_flip(void* const* p, int esz, const uint8_t* solid, long n, const double* stagnant, hipStream_t s) {
    unsigned nb = (unsigned)((n + 255) / 256);
    if (esz == 8) ibm_flip_kernel<double><<<nb, 256, 0, s>>>((double*)p[0], (double*)p[1], (double*)p[2], (double*)p[3], solid, n, stagnant);
    else ibm_flip_kernel<float><<<nb, 256, 0, s>>>((float*)p[0], (float*)p[1], (float*)p[2], (float*)p[3], solid, n, stagnant);
    return hipGetLastError();
}
}  // namespace efvm
