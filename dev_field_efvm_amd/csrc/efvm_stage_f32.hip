// efvm_stage_f32.hip -- FP32 production stage kernel (multiply-adds fused within one expression).
//
// Same wave-strip decomposition and boundary semantics as the FP64 parity
// kernel (efvm_stage_f64.hip); the numerics are the 2-D schemes of efvm_fast.h.
// The z faces of the one-cell-thick mesh are dropped: their two contributions
// cancel exactly in exact arithmetic.  Per cell: rhs = (F_w - F_e) A_x/V +
// (G_s - G_n) A_y/V (+ viscous, body force, solid zero), then the RK combine.
// Stage 3 fuses the CFL scan: lambda of the new state per cell, max-reduced in
// FP64 across the wave and the grid.
//
// gfx950 specifics:
//  * lane shifts use DPP wave_shr:1 / wave_shl:1 (one VALU move, no LDS);
//  * the column of each lane (and its ghost/periodic mapping) is fixed for
//    the whole strip; the row index is wave-uniform, so the per-row boundary
//    logic is a scalar branch;
//  * rows are prefetched two iterations ahead and U0 one iteration ahead,
//    so HBM latency overlaps the flux arithmetic;
//  * rarely used features (viscous terms, immersed body, inflow/outlet
//    boundaries needing the EOS) are template features, so the common
//    periodic/reflecting case carries none of their registers.
#include "efvm_stage.h"
#include "efvm_fast.h"
#include "efvm_cfl_f32.h"

namespace efvm {

using F4 = fast::C4<float>;
using fast::Prep;

enum { FEAT_VISC = 1, FEAT_IBM = 2, FEAT_CBC = 4, FEAT_BODY = 8 };

// lane i receives the value of lane i-1 (from_left) / lane i+1 (from_right); the
// lane with no source (0, resp. 63) receives 0 (bound_ctrl), so the move needs no
// initialised destination
__device__ __forceinline__ float from_left(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ int from_left_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int from_right_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ F4 from_left4(const F4& v) { return {from_left(v.r), from_left(v.ru), from_left(v.rv), from_left(v.re)}; }
__device__ __forceinline__ F4 from_right4(const F4& v) { return {from_right(v.r), from_right(v.ru), from_right(v.rv), from_right(v.re)}; }
__device__ __forceinline__ Prep from_leftP(const Prep& v) {
    Prep o;
    o.r = from_left(v.r); o.ru = from_left(v.ru); o.rv = from_left(v.rv); o.re = from_left(v.re);
    o.u = from_left(v.u); o.v = from_left(v.v); o.p = from_left(v.p); o.q = from_left(v.q);
    o.a = from_left(v.a); o.ia = from_left(v.ia); o.ir = from_left(v.ir);
    return o;
}
__device__ __forceinline__ F4 sub4(const F4& a, const F4& b) { return {a.r - b.r, a.ru - b.ru, a.rv - b.rv, a.re - b.re}; }

// reflecting / copying ghost of an owner cell (SupersonicOutflow, SlipWall, Symmetry, NoSlipWall)
__device__ __forceinline__ F4 reflect(int kind, int axis, F4 v) {
    if (kind == BC_SLIP || kind == BC_SYMMETRY) {
        if (axis == 0) v.ru = -v.ru; else v.rv = -v.rv;
    } else if (kind == BC_NOSLIP) {
        v.ru = -v.ru; v.rv = -v.rv;
    }
    return v;
}

// Row r (wave-uniform) of the lane's column.  `cq` is the column to read (the
// periodic image or, for a ghost lane, the owner column); `xkind` is -1 for a
// real column, else the boundary kind of the x side the lane is a ghost of
// (selected once per lane, so no per-lane indexing of the argument block).
// A row that is not `live` is read through a zero-length descriptor: the loads are
// still issued (and return 0) so that every row step has the same memory
// operations on every path, which lets the compiler count outstanding loads
// exactly instead of waiting for all of them at the first use.
template <int FEAT>
__device__ __forceinline__ F4 load_row(const StageArgs& a, const float* const* U, int c, int cq, int xkind, int r,
                                       bool live = true) {
    int rr = r, ykind = -1;
    if (r < 0 || r >= a.ny) {
        if (a.per_y) { rr = r % a.ny; if (rr < 0) rr += a.ny; }
        else { ykind = r < 0 ? a.bc_kind[2] : a.bc_kind[3]; rr = r < 0 ? 0 : a.ny - 1; }
    }
    if ((FEAT & FEAT_CBC) && (ykind >= 0 || xkind >= 0)) {
        Cell<float> cl = load_cell(a, U, c, r);   // EOS-based ghosts, evaluated in FP64
        return {cl.r, cl.ru, cl.rv, cl.re};
    }
    const long rb = (long)rr * a.nx;
    const unsigned off = (unsigned)cq * 4u, nb = live ? (unsigned)a.nx * 4u : 0u;
    F4 v = {row_ld(U[0] + rb, nb, off), row_ld(U[1] + rb, nb, off), row_ld(U[2] + rb, nb, off), row_ld(U[3] + rb, nb, off)};
    if (ykind >= 0) v = reflect(ykind, 1, v);
    if (xkind >= 0) v = reflect(xkind, 0, v);
    return v;
}

// Half the limited slope: the face states are u -/+ half_slope.  minmod(x, y) is the
// median of (0, x, y), and the three-argument minmod is minmod applied twice, so
//   MC / 2     = minmod(a, b, (a + b) / 4) = med3(0, med3(0, a, b), (a + b) / 4),
//   minmod / 2 = med3(0, a, b) / 2,
// the same values as 0.5 * mc(a, b) and 0.5 * minmod(a, b) (the factors 1/2 and 1/4
// are exact) in two or three VALU operations.
template <int LIM>
__device__ __forceinline__ float half_slope(float a, float b) {
    const float m = __builtin_amdgcn_fmed3f(0.f, a, b);
    // med3(0, m, (a+b)/4) written with the factors 2 and 1/2 (exact, so the same value), which
    // the unit's IEEE-off build folds into the output modifiers of the med3 and add
    if (LIM == 1) return 0.5f * __builtin_amdgcn_fmed3f(0.f, 2.f * m, 0.5f * (a + b));
    return 0.5f * m;
}
template <int LIM>
__device__ __forceinline__ F4 half_slope4(const F4& dr, const F4& dl) {
    return {half_slope<LIM>(dr.r, dl.r), half_slope<LIM>(dr.ru, dl.ru), half_slope<LIM>(dr.rv, dl.rv),
            half_slope<LIM>(dr.re, dl.re)};
}
__device__ __forceinline__ F4 axpy4(const F4& u, float h, const F4& s) {
    return {fmaf(h, s.r, u.r), fmaf(h, s.ru, u.ru), fmaf(h, s.rv, u.rv), fmaf(h, s.re, u.re)};
}

struct Prim3 { float u, v, t; };
__device__ __forceinline__ Prim3 prim3(const F4& q, float gm1, float ir_gas) {
    float iq = fast::rcp(q.r);
    float u = q.ru * iq, v = q.rv * iq;
    return {u, v, (gm1 * (q.re - 0.5f * (q.ru * u + q.rv * v))) * iq * ir_gas};
}

// One wave processes strip `s` of the case `a`; planes and stage/parity are passed
// separately so the batched kernel can reuse it.  The wave owning
// strip 0 does the per-case clock bookkeeping.
//
// Row loop: a prologue builds the y "north" state of row j0-1, then one step per
// row jc = j0 .. j1 computes the y face jc-1/2, the x faces of row jc (jc < j1)
// and writes row jc-1 (jc > j0).  The steady state is unrolled four times with
// the register roles rotated (row window of 4, ping-pong for the carried face
// data), so advancing a row costs no register moves.
template <int FLUX, int LIM, int FEAT>   // LIM = LIM_NONE: first order
__device__ __forceinline__ void stage_f32_strip(const StageArgs& a, const int s, const float* const* U,
                                                const float* const* U0, float* const* O, const int stage,
                                                const int parity) {
    constexpr bool MUSCL = LIM != LIM_NONE;
    constexpr bool VISC = (FEAT & FEAT_VISC) != 0;
    constexpr bool IBM = (FEAT & FEAT_IBM) != 0;
    const int lane = lane_id();
    const int nx = a.nx, ny = a.ny;
    const int sxi = s % a.n_strips_x, syi = s / a.n_strips_x;
    const int c = sxi * STRIP_OUT - 2 + lane;
    const int j0 = syi * a.strip_h;
    const int j1 = min(j0 + a.strip_h, ny);
    const float g = (float)a.gamma, gm1 = (float)(a.gamma - 1.0);
    const float gg = (float)(a.gamma / (a.gamma - 1.0));   // gamma / (gamma - 1): enthalpy per pressure
    const double dtd = a.fixed_dt > 0.0 ? a.fixed_dt : a.cfl / a.clock->lam[parity];
    const float dt = (float)dtd;
    const bool fuse = a.fuse_cfl && (a.euler || stage == 2);
    if (s == 0 && lane == 0 && stage == 0) { a.clock->dt = dtd; a.clock->lam[parity ^ 1] = 0.0; }
    float ca, cb;   // out = ca*u0 + cb*(u + dt*rhs)
    if (a.euler || stage == 0) { ca = 0.f; cb = 1.f; }
    else if (stage == 1) { ca = 0.75f; cb = 0.25f; }
    else { ca = 1.f / 3.f; cb = 2.f / 3.f; }
    const float dtix = dt * (float)a.inv_x, dtiy = dt * (float)a.inv_y;
    const float ir_gas = (float)(1.0 / a.r_gas);
    const float kap_fac = (float)(a.gamma * a.r_gas / (a.gamma - 1.0) / 0.72);
    const float gxdt = (float)a.gx * dt;

    // ---- per-lane column mapping (fixed for the strip) ----
    const bool col_ghost = (c < 0 || c >= nx);
    int cq = c, xkind = -1;
    if (col_ghost) {
        if (a.per_x) { cq = c % nx; if (cq < 0) cq += nx; }
        else { xkind = c < 0 ? a.bc_kind[0] : a.bc_kind[1]; cq = c < 0 ? 0 : nx - 1; }
    }
    const bool seam_rec = a.seam_mode == 1;
    const bool xW_rec = MUSCL && (col_ghost ? (a.per_x && seam_rec) : !(c == 0 && !a.per_x));
    const bool xE_rec = MUSCL && (col_ghost ? (a.per_x && seam_rec) : !(c == nx - 1 && !a.per_x));
    // face state = u + h * half_slope with h = -1 (west/south), +1 (east/north) or 0
    // where the reference keeps that side first order; fma(0, s, u) == u for finite s
    const float hW = xW_rec ? -1.f : 0.f, hE = xE_rec ? 1.f : 0.f;
    const bool out_lane = lane >= 2 && lane < WAVE - 2 && c < nx;
    // U0 byte offset of the lane's column: a lane left of the mesh (c < 0) gets an offset
    // far past any row (as the halo stores do), one right of it is past the row already
    const unsigned u0off = c >= 0 ? (unsigned)c * 4u : 0x80000000u;
    auto solid_at = [&](int rr) -> int {
        if (!IBM || col_ghost || rr < 0 || rr >= ny) return 0;
        return a.solid[(long)rr * nx + c];
    };

    float lam_max = 0.f;
    int bad = 0;

    // y reconstruction of row jc: south (lo) and north (hi) face states
    auto y_states = [&](int jc, const F4& uc, const F4& fwd, const F4& bwd, Prep& lo_p, Prep& hi_p) {
        if (MUSCL) {
            const bool row_ghost = (jc < 0 || jc >= ny);
            const bool yS_rec = row_ghost ? (a.per_y && seam_rec) : !(jc == 0 && !a.per_y);
            const bool yN_rec = row_ghost ? (a.per_y && seam_rec) : !(jc == ny - 1 && !a.per_y);
            // row flags are wave-uniform: keep them (and the factors) in scalar registers
            const float hS = yS_rec ? -1.f : 0.f;
            const float hN = yN_rec ? 1.f : 0.f;
            const F4 sy = half_slope4<LIM>(fwd, bwd);
            lo_p = fast::prep(axpy4(uc, hS, sy), g, gm1);
            hi_p = fast::prep(axpy4(uc, hN, sy), g, gm1);
        } else {
            lo_p = fast::prep(uc, g, gm1);
            hi_p = lo_p;
        }
    };

    // One row step.  In: um,uc,up = U(jc-1..jc+1), un2 = U(jc+2) (in flight), din = U(jc)-U(jc-1),
    // hin = north state of row jc-1, gin = G(jc-3/2), xin = x divergence of row jc-1,
    // vin = viscous terms of row jc-1.  Out: dout, hout, gout = G(jc-1/2), xout, vout.
    // Solid flags: s_m, s_c, s_p of rows jc-1, jc, jc+1, s_n of jc+2.  Reloads um with U(jc+3).
    auto step = [&](int jc, F4& um, const F4& uc, const F4& up, F4& din, F4& dout, Prep& hin, Prep& hout,
                    F4& gin, F4& gout, F4& xin, F4& xout, float* vin, float* vout, int& s_m, int s_c, int s_p,
                    const F4& u0in, F4& u0out) {
        const bool full = jc < j1;
        const bool out = jc > j0 && jc <= j1;   // a step past j1 writes nothing
        const long rb = (long)(jc - 1) * nx;    // wave-uniform row base
        const unsigned coff = (unsigned)c * 4u;  // meaningful on output lanes only
        const unsigned rowb = (unsigned)nx * 4u;
        const F4 u0 = u0in;
        // U0 of row jc (the next step's output row): loaded a whole row step ahead of its use.
        // The loads are issued on every step and lane (a zero-length descriptor when U0 is not
        // needed; a lane outside the mesh reads 0 from past the row), so that the memory operations
        // of a step do not depend on the path and the loads in flight are counted exactly.
        {
            const long rn = (long)jc * nx;
            const unsigned u0b = (full && ca != 0.f) ? rowb : 0u;
            u0out = {row_ld(U0[0] + rn, u0b, u0off), row_ld(U0[1] + rn, u0b, u0off), row_ld(U0[2] + rn, u0b, u0off),
                     row_ld(U0[3] + rn, u0b, u0off)};
        }
        dout = sub4(up, uc);
        Prep lo_p;
        y_states(jc, uc, dout, din, lo_p, hout);
        gout = fast::flux<FLUX, 1>(hin, lo_p, gg);
        if (full) {
            Prep W_p, E_p;
            if (MUSCL) {
                const F4 ul = from_left4(uc), ur = from_right4(uc);
                const F4 sx = half_slope4<LIM>(sub4(ur, uc), sub4(uc, ul));
                W_p = fast::prep(axpy4(uc, hW, sx), g, gm1);
                E_p = fast::prep(axpy4(uc, hE, sx), g, gm1);
            } else {
                W_p = lo_p;
                E_p = lo_p;
            }
            const Prep El = from_leftP(E_p);
            const F4 Fw = fast::flux<FLUX, 0>(El, W_p, gg);
            const F4 Fe = from_right4(Fw);
            xout = sub4(Fw, Fe);
            if (VISC) {
                const Prim3 pc = prim3(uc, gm1, ir_gas);
                Prim3 pl{from_left(pc.u), from_left(pc.v), from_left(pc.t)};
                Prim3 pr{from_right(pc.u), from_right(pc.v), from_right(pc.t)};
                const int sl_ = from_left_i(s_c), sr_ = from_right_i(s_c);
                Prim3 pd = prim3(um, gm1, ir_gas), pu = prim3(up, gm1, ir_gas);
                vout[0] = 0.f; vout[1] = 0.f; vout[2] = 0.f;
                if (!s_c && !col_ghost) {
                    const float mu = a.visc_const ? (float)a.mu : (float)ref::sutherland((double)pc.t);
                    const float kap = mu * kap_fac;
                    if (sr_) pr = {0.f, 0.f, pc.t};
                    if (sl_) pl = {0.f, 0.f, pc.t};
                    if (IBM && s_p) pu = {0.f, 0.f, pc.t};
                    if (IBM && s_m) pd = {0.f, 0.f, pc.t};
                    const float yim = (float)a.yw[jc], yip = (float)a.yw[ny + jc], yis = (float)a.yw[2 * ny + jc];
                    const float dxi2 = (float)a.dxi2;
                    const float lapu = (pr.u - 2.f * pc.u + pl.u) * dxi2 + ((pu.u - pc.u) * yip - (pc.u - pd.u) * yim) * yis;
                    const float lapv = (pr.v - 2.f * pc.v + pl.v) * dxi2 + ((pu.v - pc.v) * yip - (pc.v - pd.v) * yim) * yis;
                    const float lapt = (pr.t - 2.f * pc.t + pl.t) * dxi2 + ((pu.t - pc.t) * yip - (pc.t - pd.t) * yim) * yis;
                    vout[0] = mu * lapu;
                    vout[1] = mu * lapv;
                    vout[2] = pc.u * vout[0] + pc.v * vout[1] + kap * lapt;
                }
            }
        }
        // The new row jc-1: computed on every step and lane and stored through a descriptor
        // that drops what is not an output (zero length on the first step, an offset past
        // the row on halo lanes), so the stores too are the same on every path.
        {
            const bool live = out && out_lane;
            const long q = rb + c;
            float rr = xin.r * dtix + (gin.r - gout.r) * dtiy;
            float rru = xin.ru * dtix + (gin.ru - gout.ru) * dtiy;
            float rrv = xin.rv * dtix + (gin.rv - gout.rv) * dtiy;
            float rre = xin.re * dtix + (gin.re - gout.re) * dtiy;
            if (VISC) { rru += dt * vin[0]; rrv += dt * vin[1]; rre += dt * vin[2]; }
            if ((FEAT & FEAT_BODY) && a.body_force) { rru += um.r * gxdt; rre += gxdt * um.ru; }
            if (IBM && live && a.solid[q]) { rr = 0.f; rru = 0.f; rrv = 0.f; rre = 0.f; }
            float o0, o1, o2, o3;
            if (ca != 0.f) {
                // ca u0 + cb (u + dt L) written as u0 + cb ((u - u0) + dt L): the state keeps the
                // coefficient 1 exactly (FP32 1/3 + 2/3 != 1 would scale it every step) and the
                // rounding of cb touches only the increment
                o0 = u0.r + cb * ((um.r - u0.r) + rr); o1 = u0.ru + cb * ((um.ru - u0.ru) + rru);
                o2 = u0.rv + cb * ((um.rv - u0.rv) + rrv); o3 = u0.re + cb * ((um.re - u0.re) + rre);
            } else {
                o0 = um.r + rr; o1 = um.ru + rru; o2 = um.rv + rrv; o3 = um.re + rre;
            }
            const unsigned sb = out ? rowb : 0u, so = out_lane ? coff : 0x80000000u;
            row_st(O[0] + rb, sb, so, o0); row_st(O[1] + rb, sb, so, o1); row_st(O[2] + rb, sb, so, o2);
            row_st(O[3] + rb, sb, so, o3);
            if (fuse && live) {
                float l;
                if (lam_cell_f32<VISC>(a, o0, o1, o2, o3, l)) lam_max = fmaxf(lam_max, l);
                else bad++;
            }
        }
        {
            const bool more = full && jc + 3 <= j1 + 1;
            um = load_row<FEAT>(a, U, c, cq, xkind, jc + 3, more);
            if (IBM && more) s_m = solid_at(jc + 3);
        }
    };

    // ---- prologue: rows j0-2 .. j0+2, north state of row j0-1 ----
    F4 R0 = load_row<FEAT>(a, U, c, cq, xkind, j0 - 1);
    F4 R1 = load_row<FEAT>(a, U, c, cq, xkind, j0);
    F4 R2 = load_row<FEAT>(a, U, c, cq, xkind, j0 + 1);
    F4 R3 = load_row<FEAT>(a, U, c, cq, xkind, j0 + 2);
    int S0 = 0, S1 = 0, S2 = 0, S3 = 0;
    if (IBM) { S0 = solid_at(j0 - 1); S1 = solid_at(j0); S2 = solid_at(j0 + 1); S3 = solid_at(j0 + 2); }
    F4 D0, D1;
    Prep H0, H1;
    F4 G0{0.f, 0.f, 0.f, 0.f}, G1{0.f, 0.f, 0.f, 0.f}, X0{0.f, 0.f, 0.f, 0.f}, X1{0.f, 0.f, 0.f, 0.f};
    float V0[3] = {0.f, 0.f, 0.f}, V1[3] = {0.f, 0.f, 0.f};
    F4 Q0{0.f, 0.f, 0.f, 0.f}, Q1{0.f, 0.f, 0.f, 0.f};   // U0 rows in flight (ping-pong)
    {
        const F4 rm2 = load_row<FEAT>(a, U, c, cq, xkind, j0 - 2);
        D0 = sub4(R1, R0);                 // U(j0) - U(j0-1): forward difference of row j0-1
        Prep lo_unused;
        y_states(j0 - 1, R0, D0, sub4(R0, rm2), lo_unused, H0);
    }
    // Drain the prologue's loads (vmcnt(0); expcnt and lgkmcnt left at their maxima).
    // The loop head is reached from here and from its back edge; with nothing left
    // in flight from the prologue, the waits inside the loop are the back edge's,
    // which keep two row steps of loads in flight.
    __builtin_amdgcn_s_waitcnt(0x0F70);
    // ---- steady state: rows j0 .. j1, unrolled x4 with rotated roles ----
    // The loop runs whole groups of four row steps and has no exit but its condition,
    // so the waits for the rows loaded two steps ahead are counted from the back edge
    // alone.  Steps past j1 (up to three, none when strip_rows + 1 is a multiple of 4)
    // store, load and accumulate nothing.  Which of the four compiled steps takes a row
    // depends on the strip height; the unit fuses multiply-adds only within one source
    // expression (-ffp-contract=on), so the four copies round alike and results do not
    // depend on the height.
    for (int jc = j0; jc <= j1; jc += 4) {
        step(jc, R0, R1, R2, D0, D1, H0, H1, G0, G1, X0, X1, V0, V1, S0, S1, S2, Q0, Q1);
        step(jc + 1, R1, R2, R3, D1, D0, H1, H0, G1, G0, X1, X0, V1, V0, S1, S2, S3, Q1, Q0);
        step(jc + 2, R2, R3, R0, D0, D1, H0, H1, G0, G1, X0, X1, V0, V1, S2, S3, S0, Q0, Q1);
        step(jc + 3, R3, R0, R1, D1, D0, H1, H0, G1, G0, X1, X0, V1, V0, S3, S0, S1, Q1, Q0);
    }

    if (fuse) {
        const double w = wave_max((double)lam_max);
        const long long b = wave_sum_ll((long long)bad);
        if (lane == 0) {
            if (a.lam_part) a.lam_part[s] = w;   // one slot per strip, reduced by lam_reduce_kernel
            else if (w > 0.0) atomic_max_pos_double(&a.clock->lam[parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    if (s == 0 && lane == 0 && (a.euler || stage == 2)) {
        a.clock->time += dtd;
        a.clock->step += 1;
    }
}

template <int FLUX, int LIM, int FEAT>
__global__ __launch_bounds__(256) void stage_f32_kernel(StageArgs a) {
    const int s = strip_index(a, blockDim.x / WAVE);
    if (s >= a.n_strips_x * a.n_strips_y) return;
    const float* U[4] = {(const float*)a.uin[0], (const float*)a.uin[1], (const float*)a.uin[2], (const float*)a.uin[3]};
    const float* U0[4] = {(const float*)a.u0[0], (const float*)a.u0[1], (const float*)a.u0[2], (const float*)a.u0[3]};
    float* O[4] = {(float*)a.uout[0], (float*)a.uout[1], (float*)a.uout[2], (float*)a.uout[3]};
    stage_f32_strip<FLUX, LIM, FEAT>(a, s, U, U0, O, a.stage, a.parity);
}

// ---------------------------------------------------------------------------
// Two-column strips, no boundary features, and the four planes of every set at one stride,
// so that each set is one buffer descriptor and a plane is selected by the scalar offset.
// A wave owns 128 consecutive columns, two per lane (c0 = x0 + 2 lane - 2 and c0 + 1), and
// writes the 124 columns of lanes 1..62.  Per row step a lane evaluates the x face between
// its two columns without any lane exchange, the face west of c0 from the east state of the
// left lane's second column (one DPP move per value), and the y faces of both columns: the
// lane exchanges of a cell halve, and so do the halo lanes.  Every cell's arithmetic is
// stage_f32_strip's, expression for expression (the same slopes, face states, fluxes,
// divergence, RK combine and CFL), so the results are bitwise those of the one-column kernel.
//
// GEN = false: both axes periodic, nx even; a lane loads its two columns with one 8-byte load
// per plane and stores them with one.  GEN = true: any mix of periodic, outflow, slip, no-slip
// and symmetry sides, and odd nx on a non-periodic x axis.  A ghost column takes its owner's
// value (the periodic image, or the edge column reflected as the side's condition says), a
// ghost row likewise; the lane loads the two columns at the clamped owner column of its first
// column and picks each column's value from the pair, and stores each column on its own, so
// the last pair of an odd row writes one column.
// ---------------------------------------------------------------------------
struct Pair { F4 a, b; };   // the lane's two columns
struct PrepPair { Prep a, b; };

__device__ __forceinline__ Pair sub_pair(const Pair& x, const Pair& y) { return {sub4(x.a, y.a), sub4(x.b, y.b)}; }

// ghost reflection factors (for rho_u, rho_v) of a side's condition on an axis: slip and
// symmetry flip the normal momentum, no-slip both, outflow and periodic neither
__device__ __forceinline__ float2 reflect_uv(int kind, int axis) {
    const bool normal = kind == BC_SLIP || kind == BC_SYMMETRY, both = kind == BC_NOSLIP;
    return {both || (normal && axis == 0) ? -1.f : 1.f, both || (normal && axis == 1) ? -1.f : 1.f};
}

template <int FLUX, int LIM, bool GEN>
__device__ __forceinline__ void stage_f32_pair_strip(const StageArgs& a, const int s, const float* U, const float* U0,
                                                     float* O, const int stage, const int parity) {
    constexpr bool MUSCL = LIM != LIM_NONE;
    const int lane = lane_id();
    const int nx = a.nx, ny = a.ny;
    const int sxi = s % a.n_strips_x, syi = s / a.n_strips_x;
    const int c0 = sxi * STRIP_OUT2 - 2 + 2 * lane;     // even: the pair never straddles a periodic seam
    const int j0 = syi * a.strip_h;
    const int j1 = min(j0 + a.strip_h, ny);
    const float g = (float)a.gamma, gm1 = (float)(a.gamma - 1.0);
    const float gg = (float)(a.gamma / (a.gamma - 1.0));   // gamma / (gamma - 1): enthalpy per pressure
    const double dtd = a.fixed_dt > 0.0 ? a.fixed_dt : a.cfl / a.clock->lam[parity];
    const float dt = (float)dtd;
    const bool fuse = a.fuse_cfl && (a.euler || stage == 2);
    if (s == 0 && lane == 0 && stage == 0) { a.clock->dt = dtd; a.clock->lam[parity ^ 1] = 0.0; }
    float ca, cb;
    if (a.euler || stage == 0) { ca = 0.f; cb = 1.f; }
    else if (stage == 1) { ca = 0.75f; cb = 0.25f; }
    else { ca = 1.f / 3.f; cb = 2.f / 3.f; }
    const float dtix = dt * (float)a.inv_x, dtiy = dt * (float)a.inv_y;

    const bool seam_rec = a.seam_mode == 1;
    const bool ga = c0 < 0 || c0 >= nx, gb = c0 + 1 < 0 || c0 + 1 >= nx;
    const bool out_lane = lane >= 1 && lane <= WAVE - 2 && c0 < nx;
    // columns: the load offset (two columns from there) and, for GEN, which of the two each of
    // the lane's columns takes and the reflection of a ghost column
    unsigned ld_off;
    bool sel_a = false, sel_b = true;
    float2 xa = {1.f, 1.f}, xb = {1.f, 1.f};
    float hWa, hEa, hWb, hEb;
    if (!GEN) {
        // the periodic image for loads; columns outside the mesh are seam ghosts, which seam
        // mode 0 leaves unreconstructed (as stage_f32_strip does)
        int cq0 = c0 % nx;
        if (cq0 < 0) cq0 += nx;
        ld_off = (unsigned)cq0 * 4u;
        hWa = MUSCL && (!ga || seam_rec) ? -1.f : 0.f; hEa = MUSCL && (!ga || seam_rec) ? 1.f : 0.f;
        hWb = MUSCL && (!gb || seam_rec) ? -1.f : 0.f; hEb = MUSCL && (!gb || seam_rec) ? 1.f : 0.f;
    } else {
        auto owner = [&](int c) {
            if (a.per_x) { int q = c % nx; return q < 0 ? q + nx : q; }
            return c < 0 ? 0 : (c >= nx ? nx - 1 : c);
        };
        const int oa = owner(c0), ob = owner(c0 + 1);
        const int lc = a.per_x ? oa : min(oa, nx - 2);
        sel_a = oa != lc;
        sel_b = ob != lc;
        ld_off = (unsigned)lc * 4u;
        if (!a.per_x) {
            const float2 lo = reflect_uv(a.bc_kind[0], 0), hi = reflect_uv(a.bc_kind[1], 0), one = {1.f, 1.f};
            xa = ga ? (c0 < 0 ? lo : hi) : one;
            xb = gb ? (c0 + 1 < 0 ? lo : hi) : one;
        }
        // face state = u + h * half_slope: a ghost column is reconstructed only across a
        // periodic seam in seam mode 1, and a boundary side of an edge column not at all
        auto hw = [&](int c, bool gh) { return MUSCL && (gh ? (a.per_x && seam_rec) : !(c == 0 && !a.per_x)) ? -1.f : 0.f; };
        auto he = [&](int c, bool gh) { return MUSCL && (gh ? (a.per_x && seam_rec) : !(c == nx - 1 && !a.per_x)) ? 1.f : 0.f; };
        hWa = hw(c0, ga); hEa = he(c0, ga); hWb = hw(c0 + 1, gb); hEb = he(c0 + 1, gb);
    }
    // stores and U0 loads of lanes that write nothing take an offset of 2^31, past the set (under
    // 2^31 bytes) with or without the scalar offset added, so the descriptor drops them
    const unsigned st_off = out_lane ? (unsigned)c0 * 4u : 0x80000000u;
    const unsigned st_off_b = out_lane && !gb ? (unsigned)(c0 + 1) * 4u : 0x80000000u;
    const unsigned u0_off = GEN ? (out_lane ? ld_off : 0x80000000u) : st_off;
    const unsigned ps = a.plane_stride, set_bytes = 4u * a.plane_stride, rowb = (unsigned)nx * 4u;

    // rows: one wrap suffices (ny >= 4 and rows span j0 - 2 .. j1 + 3); GEN: a non-periodic
    // ghost row is its edge row reflected as the side's condition says (wave-uniform factors)
    struct RowRef { unsigned off; float2 f; };
    auto row_of = [&](int r) -> RowRef {
        float2 f = {1.f, 1.f};
        if (r < 0 || r >= ny) {
            if (!GEN || a.per_y) r = r < 0 ? r + ny : r - ny;
            else {
                f = reflect_uv(r < 0 ? a.bc_kind[2] : a.bc_kind[3], 1);
                r = r < 0 ? 0 : ny - 1;
            }
        }
        return {(unsigned)r * rowb, f};
    };
    // a row of the two columns from a set; `live` false issues the loads through a zero-length
    // descriptor (they return 0), so every row step has the same memory operations
    auto ld_pair = [&](const float* set, unsigned roff, unsigned voff, bool live) -> Pair {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)set, 0, live ? (int)set_bytes : 0, BUF_DWORD3);
        Pair p;
        float v[4][2];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const float2 x = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, roff + (unsigned)m * ps, 0));
            v[m][0] = x.x;
            v[m][1] = x.y;
        }
        if (GEN) {
            p.a = {sel_a ? v[0][1] : v[0][0], sel_a ? v[1][1] : v[1][0], sel_a ? v[2][1] : v[2][0], sel_a ? v[3][1] : v[3][0]};
            p.b = {sel_b ? v[0][1] : v[0][0], sel_b ? v[1][1] : v[1][0], sel_b ? v[2][1] : v[2][0], sel_b ? v[3][1] : v[3][0]};
        } else {
            p.a = {v[0][0], v[1][0], v[2][0], v[3][0]};
            p.b = {v[0][1], v[1][1], v[2][1], v[3][1]};
        }
        return p;
    };
    auto load_row = [&](int r, bool live) -> Pair {
        const RowRef rw = row_of(r);
        Pair p = ld_pair(U, rw.off, ld_off, live);
        if (GEN) {
            // the y and x reflections of stage_f32_strip's loader (each a sign flip)
            p.a.ru *= rw.f.x * xa.x; p.a.rv *= rw.f.y * xa.y;
            p.b.ru *= rw.f.x * xb.x; p.b.rv *= rw.f.y * xb.y;
        }
        return p;
    };

    float lam_max = 0.f;
    int bad = 0;

    // y reconstruction of row jc for both columns
    auto y_states = [&](int jc, const Pair& uc, const Pair& fwd, const Pair& bwd, PrepPair& lo_p, PrepPair& hi_p) {
        if (MUSCL) {
            const bool row_ghost = (jc < 0 || jc >= ny);
            float hS, hN;
            if (!GEN) {
                const bool rec = !row_ghost || seam_rec;
                hS = rec ? -1.f : 0.f;
                hN = rec ? 1.f : 0.f;
            } else {
                const bool yS = row_ghost ? (a.per_y && seam_rec) : !(jc == 0 && !a.per_y);
                const bool yN = row_ghost ? (a.per_y && seam_rec) : !(jc == ny - 1 && !a.per_y);
                hS = yS ? -1.f : 0.f;
                hN = yN ? 1.f : 0.f;
            }
            const F4 sa = half_slope4<LIM>(fwd.a, bwd.a), sb = half_slope4<LIM>(fwd.b, bwd.b);
            lo_p.a = fast::prep(axpy4(uc.a, hS, sa), g, gm1);
            hi_p.a = fast::prep(axpy4(uc.a, hN, sa), g, gm1);
            lo_p.b = fast::prep(axpy4(uc.b, hS, sb), g, gm1);
            hi_p.b = fast::prep(axpy4(uc.b, hN, sb), g, gm1);
        } else {
            lo_p.a = fast::prep(uc.a, g, gm1);
            lo_p.b = fast::prep(uc.b, g, gm1);
            hi_p = lo_p;
        }
    };

    // one row step (the roles of stage_f32_strip's step, for the pair)
    auto step = [&](int jc, Pair& um, const Pair& uc, const Pair& up, Pair& din, Pair& dout, PrepPair& hin,
                    PrepPair& hout, Pair& gin, Pair& gout, Pair& xin, Pair& xout, const Pair& u0in, Pair& u0out) {
        const bool full = jc < j1;
        const bool out = jc > j0 && jc <= j1;
        u0out = ld_pair(U0, (unsigned)jc * rowb, u0_off, full && ca != 0.f);
        dout = sub_pair(up, uc);
        PrepPair lo_p;
        y_states(jc, uc, dout, din, lo_p, hout);
        gout.a = fast::flux<FLUX, 1>(hin.a, lo_p.a, gg);
        gout.b = fast::flux<FLUX, 1>(hin.b, lo_p.b, gg);
        if (full) {
            Prep Wa, Ea, Wb, Eb;
            if (MUSCL) {
                const F4 uL = from_left4(uc.b), uR = from_right4(uc.a);   // columns c0 - 1 and c0 + 2
                const F4 d01 = sub4(uc.b, uc.a);
                const F4 sa = half_slope4<LIM>(d01, sub4(uc.a, uL));
                const F4 sb = half_slope4<LIM>(sub4(uR, uc.b), d01);
                Wa = fast::prep(axpy4(uc.a, hWa, sa), g, gm1);
                Ea = fast::prep(axpy4(uc.a, hEa, sa), g, gm1);
                Wb = fast::prep(axpy4(uc.b, hWb, sb), g, gm1);
                Eb = fast::prep(axpy4(uc.b, hEb, sb), g, gm1);
            } else {
                Wa = lo_p.a; Ea = lo_p.a; Wb = lo_p.b; Eb = lo_p.b;
            }
            const F4 Fmid = fast::flux<FLUX, 0>(Ea, Wb, gg);             // the face between the pair
            const F4 Fwa = fast::flux<FLUX, 0>(from_leftP(Eb), Wa, gg);  // the face west of c0
            const F4 Fe = from_right4(Fwa);                          // the face east of c0 + 1
            xout.a = sub4(Fwa, Fmid);
            xout.b = sub4(Fmid, Fe);
        }
        {
            const bool live = out && out_lane;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)O, 0, out ? (int)set_bytes : 0, BUF_DWORD3);
            const unsigned roff = (unsigned)(jc - 1) * rowb;
            float o[4][2];
            const F4* X[2] = {&xin.a, &xin.b};
            const F4* G0[2] = {&gin.a, &gin.b};
            const F4* G1[2] = {&gout.a, &gout.b};
            const F4* UM[2] = {&um.a, &um.b};
            const F4* Q[2] = {&u0in.a, &u0in.b};
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const F4 &xi = *X[k], &gi = *G0[k], &go = *G1[k], &uu = *UM[k], &u0 = *Q[k];
                const float rr = xi.r * dtix + (gi.r - go.r) * dtiy;
                const float rru = xi.ru * dtix + (gi.ru - go.ru) * dtiy;
                const float rrv = xi.rv * dtix + (gi.rv - go.rv) * dtiy;
                const float rre = xi.re * dtix + (gi.re - go.re) * dtiy;
                if (ca != 0.f) {
                    o[0][k] = u0.r + cb * ((uu.r - u0.r) + rr); o[1][k] = u0.ru + cb * ((uu.ru - u0.ru) + rru);
                    o[2][k] = u0.rv + cb * ((uu.rv - u0.rv) + rrv); o[3][k] = u0.re + cb * ((uu.re - u0.re) + rre);
                } else {
                    o[0][k] = uu.r + rr; o[1][k] = uu.ru + rru; o[2][k] = uu.rv + rrv; o[3][k] = uu.re + rre;
                }
                // the second column of an odd row's last pair is a ghost: no CFL contribution
                if (fuse && live && (k == 0 || !GEN || !gb)) {
                    float l;
                    if (lam_cell_f32<false>(a, o[0][k], o[1][k], o[2][k], o[3][k], l)) lam_max = fmaxf(lam_max, l);
                    else bad++;
                }
            }
            if (!GEN) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    typedef int v2i __attribute__((ext_vector_type(2)));
                    const v2i w = {__builtin_bit_cast(int, o[m][0]), __builtin_bit_cast(int, o[m][1])};
                    __builtin_amdgcn_raw_buffer_store_b64(w, rs, st_off, roff + (unsigned)m * ps, 0);
                }
            } else {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, o[m][0]), rs, st_off, roff + (unsigned)m * ps, 0);
                    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, o[m][1]), rs, st_off_b, roff + (unsigned)m * ps, 0);
                }
            }
        }
        um = load_row(jc + 3, full && jc + 3 <= j1 + 1);
    };

    // prologue: rows j0-2 .. j0+2, north states of row j0-1
    Pair R0 = load_row(j0 - 1, true), R1 = load_row(j0, true), R2 = load_row(j0 + 1, true), R3 = load_row(j0 + 2, true);
    Pair D0, D1;
    PrepPair H0, H1;
    const F4 z4{0.f, 0.f, 0.f, 0.f};
    Pair G0{z4, z4}, G1{z4, z4}, X0{z4, z4}, X1{z4, z4}, Q0{z4, z4}, Q1{z4, z4};
    {
        const Pair rm2 = load_row(j0 - 2, true);
        D0 = sub_pair(R1, R0);
        PrepPair lo_unused;
        y_states(j0 - 1, R0, D0, sub_pair(R0, rm2), lo_unused, H0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    for (int jc = j0; jc <= j1; jc += 4) {
        step(jc, R0, R1, R2, D0, D1, H0, H1, G0, G1, X0, X1, Q0, Q1);
        step(jc + 1, R1, R2, R3, D1, D0, H1, H0, G1, G0, X1, X0, Q1, Q0);
        step(jc + 2, R2, R3, R0, D0, D1, H0, H1, G0, G1, X0, X1, Q0, Q1);
        step(jc + 3, R3, R0, R1, D1, D0, H1, H0, G1, G0, X1, X0, Q1, Q0);
    }

    if (fuse) {
        const double w = wave_max((double)lam_max);
        const long long b = wave_sum_ll((long long)bad);
        if (lane == 0) {
            if (a.lam_part) a.lam_part[s] = w;
            else if (w > 0.0) atomic_max_pos_double(&a.clock->lam[parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    if (s == 0 && lane == 0 && (a.euler || stage == 2)) {
        a.clock->time += dtd;
        a.clock->step += 1;
    }
}

// A strip whose 128 columns and rows j0-2 .. j1+3 all lie inside the mesh meets no boundary:
// no ghost column or row, no wrap, every face reconstructed.  The periodic form's code computes
// it exactly as the general form does, without the general form's per-row column selects and
// reflections on freshly loaded values (which cost a 2048^2 periodic case 42% in the general
// form).  The choice is wave-uniform.
__device__ __forceinline__ bool pair_strip_interior(const StageArgs& a, int s) {
    const int sxi = s % a.n_strips_x, syi = s / a.n_strips_x;
    const int c_lo = sxi * STRIP_OUT2 - 2, c_hi = c_lo + 2 * WAVE - 1;
    const int j0 = syi * a.strip_h, j1 = min(j0 + a.strip_h, a.ny);
    return c_lo >= 0 && c_hi < a.nx && j0 >= 2 && j1 + 3 < a.ny;
}

template <int FLUX, int LIM, bool GEN>
__global__ __launch_bounds__(256) void stage_f32_pair_kernel(StageArgs a) {
    const int s = strip_index(a, blockDim.x / WAVE);
    if (s >= a.n_strips_x * a.n_strips_y) return;
    if (GEN && pair_strip_interior(a, s))
        stage_f32_pair_strip<FLUX, LIM, false>(a, s, (const float*)a.uin[0], (const float*)a.u0[0], (float*)a.uout[0],
                                               a.stage, a.parity);
    else
        stage_f32_pair_strip<FLUX, LIM, GEN>(a, s, (const float*)a.uin[0], (const float*)a.u0[0], (float*)a.uout[0],
                                             a.stage, a.parity);
}

// ---------------------------------------------------------------------------
// Stages 2 and 3 of SSP-RK3 in one pass (the periodic two-column path).
//
// A workgroup of two waves owns a strip.  Wave A (role 0) is stage 2: it reads U1 and U^n and
// finishes one row of U2 per row step, rows j0-2 .. j1+1 (the strip's rows and the two halo rows
// on each side that stage 3 needs), and puts each row in a two-slot LDS ring.  Wave B (role 1)
// is stage 3: after each row step's barrier it takes the row A just finished, two row steps
// behind A, reads U^n again (the row A read a few steps earlier: a cache hit) and stores rows
// j0 .. j1-1 of U^{n+1}, with the fused CFL.  U2 never goes to memory: per cell and step the two
// stages read 32 B and write 16 B instead of 64 and 32, and each wave carries one pipeline's
// registers, so three waves fit on a SIMD.
//
// A's results are valid on lanes 1..62 (its x stencil reaches one lane each side), so B stores
// lanes 2..61: STRIP_OUT23 = 120 columns per wave.  A computes the halo cells of U2 (ghost rows
// and columns across a periodic seam among them) as the waves that own their periodic images
// do: every cell reconstructed, and, in seam mode 0, a seam face evaluated twice, once as each
// side's owner sees it (the cell across the seam unreconstructed).  B, which only writes real
// cells, applies stage_f32_pair_strip's ghost rules.  The results are bitwise those of
// stage_f32_pair_kernel run as stages 2 and 3.
// ---------------------------------------------------------------------------
// SEAMS: seam mode 0 and the strip holds a periodic seam (an x seam lane, or the first or last
// row band): the seam faces are evaluated twice and B applies the ghost rules.  Elsewhere every
// cell the wave touches is reconstructed.  FUSE: the CFL epilogue.
template <int FLUX, int LIM, bool SEAMS, bool FUSE>
__device__ __forceinline__ void stage_f32_pair23_strip(const StageArgs& a, const int s, const int role, float2* ring,
                                                       const float* U1, const float* U0, float* O, const int parity) {
    constexpr bool MUSCL = LIM != LIM_NONE;
    const int lane = lane_id();
    const int nx = a.nx, ny = a.ny;
    const int sxi = s % a.n_strips_x, syi = s / a.n_strips_x;
    const int c0 = sxi * STRIP_OUT23 - 4 + 2 * lane;    // even: the pair never straddles a periodic seam
    const int j0 = syi * a.strip_h;
    const int j1 = min(j0 + a.strip_h, ny);
    const int a0 = j0 - 2, a1 = j1 + 2;                 // A finishes U2 rows a0 .. a1-1
    const float g = (float)a.gamma, gm1 = (float)(a.gamma - 1.0);
    const float gg = (float)(a.gamma / (a.gamma - 1.0));   // gamma / (gamma - 1): enthalpy per pressure
    const double dtd = a.fixed_dt > 0.0 ? a.fixed_dt : a.cfl / a.clock->lam[parity];
    const float dt = (float)dtd;
    const float dtix = dt * (float)a.inv_x, dtiy = dt * (float)a.inv_y;
    const float cbA = 0.25f, cbB = 2.f / 3.f;          // the stage-2 and stage-3 increments
    const bool seam_rec = a.seam_mode == 1;
    int cq0 = c0 % nx;
    if (cq0 < 0) cq0 += nx;
    const unsigned ld_off = (unsigned)cq0 * 4u;
    // A: every column reconstructed; B: a ghost column only across the seam in seam mode 1
    const float hW = MUSCL ? -1.f : 0.f, hE = MUSCL ? 1.f : 0.f;
    const bool ga = c0 < 0 || c0 >= nx, gb = c0 + 1 < 0 || c0 + 1 >= nx;
    const float hWa = MUSCL && (!ga || seam_rec) ? -1.f : 0.f, hEa = MUSCL && (!ga || seam_rec) ? 1.f : 0.f;
    const float hWb = MUSCL && (!gb || seam_rec) ? -1.f : 0.f, hEb = MUSCL && (!gb || seam_rec) ? 1.f : 0.f;
    // seam mode 0: lanes whose first column lies on the x seam (its west face is the seam face),
    // and whether the wave has any (wave-uniform)
    const bool xseam_lane = SEAMS && !seam_rec && MUSCL && cq0 == 0;
    const bool xseam_wave = SEAMS && __builtin_amdgcn_readfirstlane((int)(__ballot(xseam_lane) != 0)) != 0;
    const bool out_lane = lane >= 2 && lane <= WAVE - 3 && c0 < nx;
    const unsigned st_off = out_lane ? (unsigned)c0 * 4u : 0x80000000u;
    const unsigned ps = a.plane_stride, set_bytes = 4u * a.plane_stride, rowb = (unsigned)nx * 4u;

    // row byte offset with one periodic wrap (ny >= 8; rows span j0 - 4 .. j1 + 5)
    auto row_off = [&](int r) -> unsigned { return (unsigned)(r < 0 ? r + ny : (r >= ny ? r - ny : r)) * rowb; };
    auto ld_pair = [&](const float* set, unsigned roff, unsigned voff, bool live) -> Pair {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)set, 0, live ? (int)set_bytes : 0, BUF_DWORD3);
        Pair p;
        float v[4][2];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const float2 x = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, roff + (unsigned)m * ps, 0));
            v[m][0] = x.x;
            v[m][1] = x.y;
        }
        p.a = {v[0][0], v[1][0], v[2][0], v[3][0]};
        p.b = {v[0][1], v[1][1], v[2][1], v[3][1]};
        return p;
    };
    auto load_row = [&](int r, bool live) -> Pair { return ld_pair(U1, row_off(r), ld_off, live); };

    float lam_max = 0.f;
    int bad = 0;

    // y reconstruction of one row for both columns (rec: reconstructed, else the cell value)
    auto y_states = [&](bool rec, const Pair& uc, const Pair& fwd, const Pair& bwd, PrepPair& lo_p, PrepPair& hi_p) {
        if (MUSCL) {
            const float hS = rec ? -1.f : 0.f, hN = rec ? 1.f : 0.f;
            const F4 sa = half_slope4<LIM>(fwd.a, bwd.a), sb = half_slope4<LIM>(fwd.b, bwd.b);
            lo_p.a = fast::prep(axpy4(uc.a, hS, sa), g, gm1);
            hi_p.a = fast::prep(axpy4(uc.a, hN, sa), g, gm1);
            lo_p.b = fast::prep(axpy4(uc.b, hS, sb), g, gm1);
            hi_p.b = fast::prep(axpy4(uc.b, hN, sb), g, gm1);
        } else {
            lo_p.a = fast::prep(uc.a, g, gm1);
            lo_p.b = fast::prep(uc.b, g, gm1);
            hi_p = lo_p;
        }
    };
    // the x faces of a row (stage_f32_pair_strip's step) with the reconstruction flags given;
    // `dual`: A's seam lanes take the seam face west of c0 as their own column's owner sees it
    // and hand the left lane the face as that lane's owner sees it
    auto x_faces = [&](bool dual, float hwa, float hea, float hwb, float heb, const Pair& uc,
                       const PrepPair& lo_p, Pair& xout) {
        Prep Wa, Ea, Wb, Eb;
        if (MUSCL) {
            const F4 uL = from_left4(uc.b), uR = from_right4(uc.a);   // columns c0 - 1 and c0 + 2
            const F4 d01 = sub4(uc.b, uc.a);
            const F4 sa = half_slope4<LIM>(d01, sub4(uc.a, uL));
            const F4 sb = half_slope4<LIM>(sub4(uR, uc.b), d01);
            Wa = fast::prep(axpy4(uc.a, hwa, sa), g, gm1);
            Ea = fast::prep(axpy4(uc.a, hea, sa), g, gm1);
            Wb = fast::prep(axpy4(uc.b, hwb, sb), g, gm1);
            Eb = fast::prep(axpy4(uc.b, heb, sb), g, gm1);
        } else {
            Wa = lo_p.a; Ea = lo_p.a; Wb = lo_p.b; Eb = lo_p.b;
        }
        const F4 Fmid = fast::flux<FLUX, 0>(Ea, Wb, gg);
        const Prep EbL = from_leftP(Eb);
        F4 Fwa = fast::flux<FLUX, 0>(EbL, Wa, gg);
        F4 Fgive = Fwa;
        if (dual) {
            const F4 Fown = fast::flux<FLUX, 0>(from_leftP(fast::prep(uc.b, g, gm1)), Wa, gg);
            const F4 Fleft = fast::flux<FLUX, 0>(EbL, fast::prep(uc.a, g, gm1), gg);
            Fwa = {fast::pick(xseam_lane, Fown.r, Fwa.r), fast::pick(xseam_lane, Fown.ru, Fwa.ru),
                   fast::pick(xseam_lane, Fown.rv, Fwa.rv), fast::pick(xseam_lane, Fown.re, Fwa.re)};
            Fgive = {fast::pick(xseam_lane, Fleft.r, Fgive.r), fast::pick(xseam_lane, Fleft.ru, Fgive.ru),
                     fast::pick(xseam_lane, Fleft.rv, Fgive.rv), fast::pick(xseam_lane, Fleft.re, Fgive.re)};
        }
        const F4 Fe = from_right4(Fgive);
        xout.a = sub4(Fwa, Fmid);
        xout.b = sub4(Fmid, Fe);
    };
    // increment-form RK combine of one column (stage_f32_pair_strip's expressions)
    auto combine = [&](float cb, const F4& xi, const F4& gi, const F4& go, const F4& uu, const F4& u0) -> F4 {
        const float rr = xi.r * dtix + (gi.r - go.r) * dtiy;
        const float rru = xi.ru * dtix + (gi.ru - go.ru) * dtiy;
        const float rrv = xi.rv * dtix + (gi.rv - go.rv) * dtiy;
        const float rre = xi.re * dtix + (gi.re - go.re) * dtiy;
        F4 o;
        o.r = u0.r + cb * ((uu.r - u0.r) + rr); o.ru = u0.ru + cb * ((uu.ru - u0.ru) + rru);
        o.rv = u0.rv + cb * ((uu.rv - u0.rv) + rrv); o.re = u0.re + cb * ((uu.re - u0.re) + rre);
        return o;
    };

    // A: row step jc (stage 2): the y face jc-1/2 and the x faces of row jc, U2 of row jc-1 into
    // `res`; U^n of row jc for the next step; U1 of row jc+3 into the slot of row jc-1.  At a y
    // seam in seam mode 0 the face below row jc is evaluated twice: row jc-1 is finished with the
    // face its owner sees (row jc unreconstructed), and row jc carries the other one.
    auto stepA = [&](int jc, Pair& um, const Pair& uc, const Pair& up, Pair& din, Pair& dout, PrepPair& hin,
                     PrepPair& hout, Pair& gin, Pair& gout, Pair& xin, Pair& xout, const Pair& u0in, Pair& u0out,
                     Pair& res) {
        const bool full = jc < a1;
        u0out = ld_pair(U0, row_off(jc), ld_off, full);
        dout = sub_pair(up, uc);
        PrepPair lo_p;
        y_states(true, uc, dout, din, lo_p, hout);
        Pair gnow;
        if (SEAMS && !seam_rec && MUSCL && jc % ny == 0) {
            const Prep ua = fast::prep(uc.a, g, gm1), ub = fast::prep(uc.b, g, gm1);
            gnow.a = fast::flux<FLUX, 1>(hin.a, ua, gg);
            gnow.b = fast::flux<FLUX, 1>(hin.b, ub, gg);
            gout.a = fast::flux<FLUX, 1>(fast::prep(um.a, g, gm1), lo_p.a, gg);
            gout.b = fast::flux<FLUX, 1>(fast::prep(um.b, g, gm1), lo_p.b, gg);
        } else {
            gout.a = fast::flux<FLUX, 1>(hin.a, lo_p.a, gg);
            gout.b = fast::flux<FLUX, 1>(hin.b, lo_p.b, gg);
            gnow = gout;
        }
        if (full) x_faces(xseam_wave, hW, hE, hW, hE, uc, lo_p, xout);
        res.a = combine(cbA, xin.a, gin.a, gnow.a, um.a, u0in.a);
        res.b = combine(cbA, xin.b, gin.b, gnow.b, um.b, u0in.b);
        um = load_row(jc + 3, full && jc + 3 <= a1 + 1);
    };
    // B: row step jb (stage 3) on U2 rows jb-1 (um), jb (uc), jb+1 (up, A's result of this step),
    // with stage_f32_pair_strip's ghost rules; steps j0-2 and j0-1 only build the carried
    // difference and north state
    auto stepB = [&](int jb, const Pair& um, const Pair& uc, const Pair& up, Pair& din, Pair& dout, PrepPair& hin,
                     PrepPair& hout, Pair& gin, Pair& gout, Pair& xin, Pair& xout, const Pair& u0in, Pair& u0out) {
        const bool full = jb >= j0 && jb < j1;
        const bool out = jb > j0 && jb <= j1;
        u0out = ld_pair(U0, (unsigned)max(jb, 0) * rowb, st_off, full);
        dout = sub_pair(up, uc);
        PrepPair lo_p;
        y_states(!SEAMS || !(jb < 0 || jb >= ny) || seam_rec, uc, dout, din, lo_p, hout);
        if (jb >= j0) {
            gout.a = fast::flux<FLUX, 1>(hin.a, lo_p.a, gg);
            gout.b = fast::flux<FLUX, 1>(hin.b, lo_p.b, gg);
        }
        if (full) {
            if (SEAMS) x_faces(false, hWa, hEa, hWb, hEb, uc, lo_p, xout);
            else x_faces(false, hW, hE, hW, hE, uc, lo_p, xout);
        }
        const bool live = out && out_lane;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)O, 0, out ? (int)set_bytes : 0, BUF_DWORD3);
        const unsigned roff = (unsigned)max(jb - 1, 0) * rowb;
        const F4 oa = combine(cbB, xin.a, gin.a, gout.a, um.a, u0in.a);
        const F4 ob = combine(cbB, xin.b, gin.b, gout.b, um.b, u0in.b);
        if (FUSE) {
            // every lane evaluates; selects keep the lanes and rows that store nothing out
            float la, lb;
            const bool oka = lam_cell_f32<false>(a, oa.r, oa.ru, oa.rv, oa.re, la);
            const bool okb = lam_cell_f32<false>(a, ob.r, ob.ru, ob.rv, ob.re, lb);
            lam_max = fmaxf(lam_max, fmaxf(fast::pick(live && oka, la, 0.f), fast::pick(live && okb, lb, 0.f)));
            bad += (int)(live && !oka) + (int)(live && !okb);
        }
        typedef int v2i __attribute__((ext_vector_type(2)));
        const float o[4][2] = {{oa.r, ob.r}, {oa.ru, ob.ru}, {oa.rv, ob.rv}, {oa.re, ob.re}};
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const v2i w = {__builtin_bit_cast(int, o[m][0]), __builtin_bit_cast(int, o[m][1])};
            __builtin_amdgcn_raw_buffer_store_b64(w, rs, st_off, roff + (unsigned)m * ps, 0);
        }
    };

    // the ring slot of the U2 row A finishes at its row step t (row t-1), plane m
    auto slot = [&](int t, int m) -> float2* { return ring + (((t - 1) & 1) * 4 + m) * WAVE + lane; };
    const F4 z4{0.f, 0.f, 0.f, 0.f};
    if (role == 0) {
        auto put = [&](int t, const Pair& r) {
            *slot(t, 0) = {r.a.r, r.b.r}; *slot(t, 1) = {r.a.ru, r.b.ru};
            *slot(t, 2) = {r.a.rv, r.b.rv}; *slot(t, 3) = {r.a.re, r.b.re};
        };
        // A's prologue: U1 rows a0-2 .. a0+2, the north state of row a0-1 (reconstructed: A's rows all are)
        Pair R0 = load_row(a0 - 1, true), R1 = load_row(a0, true), R2 = load_row(a0 + 1, true), R3 = load_row(a0 + 2, true);
        Pair D0, D1, T;
        PrepPair H0, H1;
        Pair G0{z4, z4}, G1{z4, z4}, X0{z4, z4}, X1{z4, z4}, Q0{z4, z4}, Q1{z4, z4};
        {
            const Pair rm2 = load_row(a0 - 2, true);
            D0 = sub_pair(R1, R0);
            PrepPair lo_unused;
            y_states(true, R0, D0, sub_pair(R0, rm2), lo_unused, H0);
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);
        // rows a0 (faces only) and a0+1 (finishes U2 row a0 = j0-2), then in step with B
        stepA(a0, R0, R1, R2, D0, D1, H0, H1, G0, G1, X0, X1, Q0, Q1, T); put(a0, T); __syncthreads();
        stepA(a0 + 1, R1, R2, R3, D1, D0, H1, H0, G1, G0, X1, X0, Q1, Q0, T); put(a0 + 1, T); __syncthreads();
        for (int t = j0; t <= a1; t += 4) {
            stepA(t, R2, R3, R0, D0, D1, H0, H1, G0, G1, X0, X1, Q0, Q1, T); put(t, T); __syncthreads();
            stepA(t + 1, R3, R0, R1, D1, D0, H1, H0, G1, G0, X1, X0, Q1, Q0, T); put(t + 1, T); __syncthreads();
            stepA(t + 2, R0, R1, R2, D0, D1, H0, H1, G0, G1, X0, X1, Q0, Q1, T); put(t + 2, T); __syncthreads();
            stepA(t + 3, R1, R2, R3, D1, D0, H1, H0, G1, G0, X1, X0, Q1, Q0, T); put(t + 3, T); __syncthreads();
        }
        return;
    }
    auto get = [&](int t) -> Pair {
        const float2 r = *slot(t, 0), u = *slot(t, 1), v = *slot(t, 2), e = *slot(t, 3);
        return {{r.x, u.x, v.x, e.x}, {r.y, u.y, v.y, e.y}};
    };
    // B: U2 rows in a window of four, each taken from the ring after A's barrier
    Pair S0{z4, z4}, S1, S2, S3;
    Pair E0{z4, z4}, E1;
    PrepPair K0, K1;
    Pair GB0{z4, z4}, GB1{z4, z4}, XB0{z4, z4}, XB1{z4, z4}, QB0{z4, z4}, QB1{z4, z4};
    {
        // defined values for the first warm-up step (whose y flux is not used)
        const Prep p0 = fast::prep(z4, g, gm1);
        K0 = {p0, p0};
    }
    __syncthreads();
    __syncthreads();
    S1 = get(a0 + 1);                                   // U2 row j0-2
    for (int t = j0; t <= a1; t += 4) {
        __syncthreads(); S2 = get(t);
        stepB(t - 2, S0, S1, S2, E0, E1, K0, K1, GB0, GB1, XB0, XB1, QB0, QB1);
        __syncthreads(); S3 = get(t + 1);
        stepB(t - 1, S1, S2, S3, E1, E0, K1, K0, GB1, GB0, XB1, XB0, QB1, QB0);
        __syncthreads(); S0 = get(t + 2);
        stepB(t, S2, S3, S0, E0, E1, K0, K1, GB0, GB1, XB0, XB1, QB0, QB1);
        __syncthreads(); S1 = get(t + 3);
        stepB(t + 1, S3, S0, S1, E1, E0, K1, K0, GB1, GB0, XB1, XB0, QB1, QB0);
    }

    if (FUSE) {
        const double w = wave_max((double)lam_max);
        const long long b = wave_sum_ll((long long)bad);
        if (lane == 0) {
            if (a.lam_part) a.lam_part[s] = w;
            else if (w > 0.0) atomic_max_pos_double(&a.clock->lam[parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    if (s == 0 && lane == 0) {
        a.clock->time += dtd;
        a.clock->step += 1;
    }
}

// one strip of the fused launch, in the seam-aware instance where the strip holds a periodic seam
// (seam mode 0, MUSCL)
template <int FLUX, int LIM, bool FUSE>
__device__ __forceinline__ void pair23_run(const StageArgs& a, int s, int role, float2* ring, const float* U1,
                                           const float* U0, float* O, int parity) {
    bool seams = false;
    if (a.seam_mode == 0 && LIM != LIM_NONE) {
        const int sxi = s % a.n_strips_x, j0 = (s / a.n_strips_x) * a.strip_h;
        const int cfirst = sxi * STRIP_OUT23 - 4, clast = cfirst + 2 * WAVE - 1;
        const bool xs = cfirst <= 0 || clast >= a.nx;                 // holds column 0 or column nx
        const bool ys = j0 - 2 <= 0 || min(j0 + a.strip_h, a.ny) + 2 >= a.ny;
        seams = xs || ys;
    }
    if (seams) stage_f32_pair23_strip<FLUX, LIM, true, FUSE>(a, s, role, ring, U1, U0, O, parity);
    else stage_f32_pair23_strip<FLUX, LIM, false, FUSE>(a, s, role, ring, U1, U0, O, parity);
}

template <int FLUX, int LIM, bool FUSE>
__global__ __launch_bounds__(2 * WAVE) void stage_f32_pair23_kernel(StageArgs a) {
    __shared__ float2 ring[2 * 4 * WAVE];
    const int s = __builtin_amdgcn_readfirstlane(block_index(a));   // one strip per workgroup
    const int role = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    pair23_run<FLUX, LIM, FUSE>(a, s, role, ring, (const float*)a.uin[0], (const float*)a.u0[0], (float*)a.uout[0],
                                a.parity);
}

// The batch's fused work items: one (case, strip) of a fused case per workgroup, its arguments
// in the fused geometry; U1 in set B, U^n in set A (even step parity) or C (odd), U^{n+1} into
// the other of the two.
template <int FLUX, int LIM, bool FUSE>
__global__ __launch_bounds__(2 * WAVE) void stage_f32_batch_pair23_kernel(const BatchCase* __restrict__ cases,
                                                                          const int2* __restrict__ work, int parity) {
    __shared__ float2 ring[2 * 4 * WAVE];
    const int2 wk = work[blockIdx.x];
    const int ci = __builtin_amdgcn_readfirstlane(wk.x), s = __builtin_amdgcn_readfirstlane(wk.y);
    const int role = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const BatchCase& bc = cases[ci];
    const int cur = (parity & 1) ? 2 : 0;
    pair23_run<FLUX, LIM, FUSE>(bc.a23, s, role, ring, (const float*)bc.P[1][0], (const float*)bc.P[cur][0],
                                (float*)bc.P[cur ^ 2][0], parity);
}

// ---------------------------------------------------------------------------
// Stage from precomputed face states (learned reconstruction, periodic grids).
// F holds 16 planes of primitives, plane (axis * 2 + side) * 4 + var: the left /
// right states of the east (axis 0) and north (axis 1) face of every cell, as the
// face kernels leave them.  One thread per cell: the Riemann fluxes of its four
// faces (each face is evaluated by both cells it separates, which keeps the kernel a
// single pass), divergence, increment-form RK combine, fused CFL, clock bookkeeping.
// ---------------------------------------------------------------------------
template <int FLUX>
__global__ __launch_bounds__(256) void stage_f32_faces_kernel(StageArgs a, const float* __restrict__ F) {
    const int nx = a.nx;
    const long np = (long)nx * a.ny;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int stage = a.stage, parity = a.parity;
    const float g = (float)a.gamma, gm1 = (float)(a.gamma - 1.0), igm1 = (float)(1.0 / (a.gamma - 1.0));
    const float gg = (float)(a.gamma / (a.gamma - 1.0));
    const double dtd = a.fixed_dt > 0.0 ? a.fixed_dt : a.cfl / a.clock->lam[parity];
    const float dt = (float)dtd;
    const bool fuse = a.fuse_cfl && (a.euler || stage == 2);
    if (q == 0 && stage == 0) { a.clock->dt = dtd; a.clock->lam[parity ^ 1] = 0.0; }
    if (q == 0 && a.face_count) *a.face_count = 0u;   // the face passes of this stage are done with it
    float ca, cb;
    if (a.euler || stage == 0) { ca = 0.f; cb = 1.f; }
    else if (stage == 1) { ca = 0.75f; cb = 0.25f; }
    else { ca = 1.f / 3.f; cb = 2.f / 3.f; }
    float lam_max = 0.f;
    int bad = 0;
    if (q < np) {
        const int j = (int)(q / nx), i = (int)(q - (long)j * nx);
        const long qw = i > 0 ? q - 1 : q + nx - 1;          // periodic west / south neighbours
        const long qs = j > 0 ? q - nx : q + np - nx;
        auto state = [&](int plane, long c) -> Prep {
            const float* f = F + plane * np + c;
            return fast::prep_prim(f[0], f[np], f[2 * np], f[3 * np], g, igm1);
        };
        const F4 Fe = fast::flux<FLUX, 0>(state(0, q), state(4, q), gg);
        const F4 Fw = fast::flux<FLUX, 0>(state(0, qw), state(4, qw), gg);
        const F4 Gn = fast::flux<FLUX, 1>(state(8, q), state(12, q), gg);
        const F4 Gs = fast::flux<FLUX, 1>(state(8, qs), state(12, qs), gg);
        const float dtix = dt * (float)a.inv_x, dtiy = dt * (float)a.inv_y;
        const float d[4] = {(Fw.r - Fe.r) * dtix + (Gs.r - Gn.r) * dtiy, (Fw.ru - Fe.ru) * dtix + (Gs.ru - Gn.ru) * dtiy,
                            (Fw.rv - Fe.rv) * dtix + (Gs.rv - Gn.rv) * dtiy, (Fw.re - Fe.re) * dtix + (Gs.re - Gn.re) * dtiy};
        float o[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float u = ((const float*)a.uin[v])[q];
            if (ca != 0.f) {   // increment form of the RK combine, as in stage_f32_strip
                const float a0 = ((const float*)a.u0[v])[q];
                o[v] = a0 + cb * ((u - a0) + d[v]);
            } else {
                o[v] = u + d[v];
            }
            ((float*)a.uout[v])[q] = o[v];
        }
        if (fuse) {
            float l;
            if (lam_cell_f32<false>(a, o[0], o[1], o[2], o[3], l)) lam_max = l;
            else bad = 1;
        }
    }
    if (fuse) {
        const double w = wave_max((double)lam_max);
        const long long b = wave_sum_ll((long long)bad);
        if (lane_id() == 0) {
            if (a.lam_part) a.lam_part[q >> 6] = w;   // one slot per wave, reduced by lam_reduce_kernel
            else if (w > 0.0) atomic_max_pos_double(&a.clock->lam[parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    if (q == 0 && (a.euler || stage == 2)) {
        a.clock->time += dtd;
        a.clock->step += 1;
    }
}

// ---------------------------------------------------------------------------
// Batched stage: hundreds of independent cases of uneven size in one launch.  A
// host-built work table lists (case, strip) pairs, largest strips first.  By default
// the grid has one wave per pair and the dispatcher balances them; a smaller grid
// strides through the table (measured up to 13% slower on 512 cases).  Every case keeps its own
// DevClock (per-case dt from its own CFL maximum) and its own plane sets
// A, B, C; all cases advance in lock-step stage by stage.
// ---------------------------------------------------------------------------
template <int FLUX, int LIM, int FEAT>
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
        stage_f32_strip<FLUX, LIM, FEAT>(bc.a, wk.y, U, U0, O, stage, parity);
    }
}

// The batch's two-column work items: cases the general two-column kernel takes (stock
// copy / reflection / periodic sides, no features), through the same strip code.
template <int FLUX, int LIM>
__global__ __launch_bounds__(256) void stage_f32_batch_pair_kernel(const BatchCase* __restrict__ cases,
                                                                   const int2* __restrict__ work, int n_work,
                                                                   int stage, int parity) {
    const int wpb = blockDim.x / WAVE;
    const int nw = gridDim.x * wpb;
    for (int w = blockIdx.x * wpb + (int)(threadIdx.x / WAVE); w < n_work; w += nw) {
        const int wu = __builtin_amdgcn_readfirstlane(w);
        const int2 wk = work[wu];
        const BatchCase& bc = cases[wk.x];
        // a fused case runs only stage 1 here, from the set holding U^n at this parity
        const int cur = bc.fused && (parity & 1) ? 2 : 0;
        const int in = stage == 0 ? cur : (stage == 1 ? 1 : 2);
        const int out = stage == 0 ? 1 : (stage == 1 ? 2 : 0);
        // the periodic form for periodic cases (as their single contexts run) and for strips
        // that meet no boundary; the general form for the others
        if (bc.a.pair_path == 1 || pair_strip_interior(bc.a, wk.y))
            stage_f32_pair_strip<FLUX, LIM, false>(bc.a, wk.y, (const float*)bc.P[in][0], (const float*)bc.P[cur][0],
                                                   (float*)bc.P[out][0], stage, parity);
        else
            stage_f32_pair_strip<FLUX, LIM, true>(bc.a, wk.y, (const float*)bc.P[in][0], (const float*)bc.P[cur][0],
                                                  (float*)bc.P[out][0], stage, parity);
    }
}

namespace {
int features_of(const StageArgs& a) {
    int f = 0;
    if (a.viscous) f |= FEAT_VISC;
    if (a.solid) f |= FEAT_IBM;
    for (int s = 0; s < 4; s++)
        if (a.bc_kind[s] == BC_SUP_INFLOW || a.bc_kind[s] == BC_SUB_INFLOW || a.bc_kind[s] == BC_PRESSURE_OUTLET ||
            a.bc_kind[s] == BC_TABLE)
            f |= FEAT_CBC;
    if (a.parabolic) f |= FEAT_CBC;
    if (a.body_force) f |= FEAT_BODY;
    return f;
}
}  // namespace

// Host dispatch over (flux, limiter, features).  The common production shapes are
// specialised; any other feature mix uses the all-features kernel.
#define EFVM_F32_DISPATCH(LAUNCH)                                                           \
    do {                                                                                    \
        const int lim = a_.muscl ? a_.limiter : LIM_NONE;                                   \
        const int ft = features_of(a_);                                                     \
        const int all = FEAT_VISC | FEAT_IBM | FEAT_CBC | FEAT_BODY;                        \
        switch (a_.flux) {                                                                  \
        case FLUX_RUSANOV:                                                                  \
            if (lim == LIM_NONE) { if (ft == 0) LAUNCH(FLUX_RUSANOV, LIM_NONE, 0); else LAUNCH(FLUX_RUSANOV, LIM_NONE, all); } \
            else if (lim == LIM_MINMOD) { if (ft == 0) LAUNCH(FLUX_RUSANOV, LIM_MINMOD, 0); else LAUNCH(FLUX_RUSANOV, LIM_MINMOD, all); } \
            else { if (ft == 0) LAUNCH(FLUX_RUSANOV, LIM_MC, 0); else LAUNCH(FLUX_RUSANOV, LIM_MC, all); } \
            break;                                                                          \
        case FLUX_HLLC:                                                                     \
            if (lim == LIM_NONE) { if (ft == 0) LAUNCH(FLUX_HLLC, LIM_NONE, 0); else LAUNCH(FLUX_HLLC, LIM_NONE, all); } \
            else if (lim == LIM_MINMOD) { if (ft == 0) LAUNCH(FLUX_HLLC, LIM_MINMOD, 0); else LAUNCH(FLUX_HLLC, LIM_MINMOD, all); } \
            else { if (ft == 0) LAUNCH(FLUX_HLLC, LIM_MC, 0); else if (ft == FEAT_CBC) LAUNCH(FLUX_HLLC, LIM_MC, FEAT_CBC); else LAUNCH(FLUX_HLLC, LIM_MC, all); } \
            break;                                                                          \
        default:                                                                            \
            if (lim == LIM_NONE) { if (ft == 0) LAUNCH(FLUX_HLLC_TEXTBOOK, LIM_NONE, 0); else LAUNCH(FLUX_HLLC_TEXTBOOK, LIM_NONE, all); } \
            else if (lim == LIM_MINMOD) { if (ft == 0) LAUNCH(FLUX_HLLC_TEXTBOOK, LIM_MINMOD, 0); else LAUNCH(FLUX_HLLC_TEXTBOOK, LIM_MINMOD, all); } \
            else { if (ft == 0) LAUNCH(FLUX_HLLC_TEXTBOOK, LIM_MC, 0); else LAUNCH(FLUX_HLLC_TEXTBOOK, LIM_MC, all); } \
            break;                                                                          \
        }                                                                                   \
    } while (0)

hipError_t launch_stage_f32(const StageArgs& a, int block, hipStream_t st) {
    const int waves = block / WAVE;
    const int nstrips = a.n_strips_x * a.n_strips_y;
    const dim3 grid((nstrips + waves - 1) / waves);
    if (a.pair_path) {
        const int lim = a.muscl ? a.limiter : LIM_NONE;
#define EFVM_PAIR_G(F, G)                                                                                   \
        do {                                                                                                \
            if (lim == LIM_MC) stage_f32_pair_kernel<F, LIM_MC, G><<<grid, block, 0, st>>>(a);              \
            else if (lim == LIM_MINMOD) stage_f32_pair_kernel<F, LIM_MINMOD, G><<<grid, block, 0, st>>>(a); \
            else stage_f32_pair_kernel<F, LIM_NONE, G><<<grid, block, 0, st>>>(a);                          \
        } while (0)
#define EFVM_PAIR(F) do { if (a.pair_path == 1) EFVM_PAIR_G(F, false); else EFVM_PAIR_G(F, true); } while (0)
        if (a.flux == FLUX_RUSANOV) EFVM_PAIR(FLUX_RUSANOV);
        else if (a.flux == FLUX_HLLC) EFVM_PAIR(FLUX_HLLC);
        else EFVM_PAIR(FLUX_HLLC_TEXTBOOK);
#undef EFVM_PAIR
#undef EFVM_PAIR_G
        return hipGetLastError();
    }
    const StageArgs& a_ = a;
#define EFVM_LAUNCH_ONE(F, L, X) stage_f32_kernel<F, L, X><<<grid, block, 0, st>>>(a)
    EFVM_F32_DISPATCH(EFVM_LAUNCH_ONE);
#undef EFVM_LAUNCH_ONE
    return hipGetLastError();
}

// Fused stages 2 and 3 (stage_f32_pair23_kernel): a.uin = U1, a.u0 = U^n, a.uout = U^{n+1};
// a.n_strips_x counts strips of STRIP_OUT23 columns; one workgroup of two waves per strip.
hipError_t launch_stage_f32_pair23(const StageArgs& a, hipStream_t st) {
    const dim3 grid(a.n_strips_x * a.n_strips_y), block(2 * WAVE);
    const int lim = a.muscl ? a.limiter : LIM_NONE;
#define EFVM_P23_C(F, C)                                                                                 \
    do {                                                                                                 \
        if (lim == LIM_MC) stage_f32_pair23_kernel<F, LIM_MC, C><<<grid, block, 0, st>>>(a);             \
        else if (lim == LIM_MINMOD) stage_f32_pair23_kernel<F, LIM_MINMOD, C><<<grid, block, 0, st>>>(a); \
        else stage_f32_pair23_kernel<F, LIM_NONE, C><<<grid, block, 0, st>>>(a);                         \
    } while (0)
#define EFVM_P23(F) do { if (a.fuse_cfl) EFVM_P23_C(F, true); else EFVM_P23_C(F, false); } while (0)
    if (a.flux == FLUX_RUSANOV) EFVM_P23(FLUX_RUSANOV);
    else if (a.flux == FLUX_HLLC) EFVM_P23(FLUX_HLLC);
    else EFVM_P23(FLUX_HLLC_TEXTBOOK);
#undef EFVM_P23
#undef EFVM_P23_C
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stage from precomputed face fluxes (learned reconstruction, periodic grids): F holds
// 8 planes, the Riemann flux through the east (planes 0..3) and north (4..7) face of
// every cell, as the face kernels leave them in out_mode 1.  One thread per cell.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void stage_f32_flux_kernel(StageArgs a, const float* __restrict__ F) {
    const int nx = a.nx;
    const long np = (long)nx * a.ny;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int stage = a.stage, parity = a.parity;
    const float g = (float)a.gamma, gm1 = (float)(a.gamma - 1.0);
    const double dtd = a.fixed_dt > 0.0 ? a.fixed_dt : a.cfl / a.clock->lam[parity];
    const float dt = (float)dtd;
    const bool fuse = a.fuse_cfl && (a.euler || stage == 2);
    if (q == 0 && stage == 0) { a.clock->dt = dtd; a.clock->lam[parity ^ 1] = 0.0; }
    if (q == 0 && a.face_count) *a.face_count = 0u;   // the face passes of this stage are done with it
    float ca, cb;
    if (a.euler || stage == 0) { ca = 0.f; cb = 1.f; }
    else if (stage == 1) { ca = 0.75f; cb = 0.25f; }
    else { ca = 1.f / 3.f; cb = 2.f / 3.f; }
    float lam_max = 0.f;
    int bad = 0;
    if (q < np) {
        const int j = (int)(q / nx), i = (int)(q - (long)j * nx);
        const long qw = i > 0 ? q - 1 : q + nx - 1;          // periodic west / south neighbours
        const long qs = j > 0 ? q - nx : q + np - nx;
        const float dtix = dt * (float)a.inv_x, dtiy = dt * (float)a.inv_y;
        float o[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float* Fx = F + v * np;
            const float* Fy = F + (4 + v) * np;
            // every multiply-add written out, so that the learned stage kernels of the other
            // unit (bitwise this kernel on the flux planes) round each cell alike
            const float d = fmaf(Fx[qw] - Fx[q], dtix, (Fy[qs] - Fy[q]) * dtiy);
            const float u = ((const float*)a.uin[v])[q];
            if (ca != 0.f) {   // increment form of the RK combine, as in stage_f32_strip
                const float a0 = ((const float*)a.u0[v])[q];
                o[v] = fmaf(cb, (u - a0) + d, a0);
            } else {
                o[v] = u + d;
            }
            ((float*)a.uout[v])[q] = o[v];
        }
        if (fuse) {
            float l;
            if (lam_cell_f32<false>(a, o[0], o[1], o[2], o[3], l)) lam_max = l;
            else bad = 1;
        }
    }
    if (fuse) {
        const double w = wave_max((double)lam_max);
        const long long b = wave_sum_ll((long long)bad);
        if (lane_id() == 0) {
            if (a.lam_part) a.lam_part[q >> 6] = w;   // one slot per wave, reduced by lam_reduce_kernel
            else if (w > 0.0) atomic_max_pos_double(&a.clock->lam[parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    if (q == 0 && (a.euler || stage == 2)) {
        a.clock->time += dtd;
        a.clock->step += 1;
    }
}

// Standalone FP32 CFL scan (first step, after an upload or a checkpoint load) with the
// same per-cell formula as the fused epilogues.
template <bool VISC>
__global__ __launch_bounds__(256) void cfl_f32_kernel(StageArgs a, const float* __restrict__ r, const float* __restrict__ ru,
                                                     const float* __restrict__ rv, const float* __restrict__ re, double* target) {
    const long n = (long)a.nx * a.ny;
    float m = 0.f;
    long long bad = 0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        float l;
        if (lam_cell_f32<VISC>(a, r[q], ru[q], rv[q], re[q], l)) m = fmaxf(m, l);
        else bad++;
    }
    const double w = wave_max((double)m);
    bad = wave_sum_ll(bad);
    if (lane_id() == 0) {
        if (w > 0.0) atomic_max_pos_double(target, w);
        if (bad) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)bad);
    }
}

__global__ void reset_target_kernel(double* t) { *t = 0.0; }

hipError_t launch_cfl_f32(const StageArgs& a, const void* const* planes, double* target, hipStream_t st) {
    const long n = (long)a.nx * a.ny;
    const long nb = (n + 255) / 256;
    const int blocks = (int)(nb < 2048 ? nb : 2048);
    const float* P[4] = {(const float*)planes[0], (const float*)planes[1], (const float*)planes[2], (const float*)planes[3]};
    reset_target_kernel<<<1, 1, 0, st>>>(target);
    if (a.viscous) cfl_f32_kernel<true><<<blocks, 256, 0, st>>>(a, P[0], P[1], P[2], P[3], target);
    else cfl_f32_kernel<false><<<blocks, 256, 0, st>>>(a, P[0], P[1], P[2], P[3], target);
    return hipGetLastError();
}

// The fused CFL's per-wave maxima -> lam[parity ^ 1] (one block).  Waves of a
// one-row-per-thread launch finish together; a max over a few thousand slots here is
// cheaper than that many atomics on one address.
__global__ __launch_bounds__(1024) void lam_reduce_kernel(const double* __restrict__ part, int n, double* target) {
    __shared__ double red[16];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, part[i]);
    m = wave_max(m);
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
        *target = fmax(*target, m);
    }
}

// The batch's per-case CFL slots -> each case's lam[parity ^ 1] (one block per case): the strips
// of the launch that ran the case's last stage (the fused launch's for a fused case).
__global__ __launch_bounds__(256) void batch_lam_fold_kernel(const BatchCase* __restrict__ cases, int parity) {
    const BatchCase& bc = cases[blockIdx.x];
    const StageArgs& a = bc.fused ? bc.a23 : bc.a;
    const int n = a.n_strips_x * a.n_strips_y;
    __shared__ double red[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, a.lam_part[i]);
    m = wave_max(m);
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) m = fmax(m, red[w]);
        double* t = &a.clock->lam[parity ^ 1];
        *t = fmax(*t, m);
    }
}

hipError_t launch_batch_lam_fold(const BatchCase* cases, int n_cases, int parity, hipStream_t st) {
    batch_lam_fold_kernel<<<n_cases, 256, 0, st>>>(cases, parity);
    return hipGetLastError();
}

hipError_t launch_lam_reduce(const double* part, int n, double* target, hipStream_t st) {
    lam_reduce_kernel<<<1, 1024, 0, st>>>(part, n, target);
    return hipGetLastError();
}

hipError_t launch_stage_f32_flux(const StageArgs& a, const float* F, hipStream_t st) {
    const long n = (long)a.nx * a.ny;
    stage_f32_flux_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a, F);
    return hipGetLastError();
}

hipError_t launch_stage_f32_faces(const StageArgs& a, const float* F, hipStream_t st) {
    const long n = (long)a.nx * a.ny;
    const unsigned grid = (unsigned)((n + 255) / 256);
    switch (a.flux) {
    case FLUX_RUSANOV: stage_f32_faces_kernel<FLUX_RUSANOV><<<grid, 256, 0, st>>>(a, F); break;
    case FLUX_HLLC: stage_f32_faces_kernel<FLUX_HLLC><<<grid, 256, 0, st>>>(a, F); break;
    default: stage_f32_faces_kernel<FLUX_HLLC_TEXTBOOK><<<grid, 256, 0, st>>>(a, F); break;
    }
    return hipGetLastError();
}

hipError_t launch_stage_f32_batch_pair(const BatchCase* cases, const StageArgs& a_, const int2* work, int n_work, int stage,
                                       int parity, int blocks, int waves, hipStream_t st) {
    const dim3 grid(blocks), block(WAVE * waves);
    const int lim = a_.muscl ? a_.limiter : LIM_NONE;
#define EFVM_BPAIR(F)                                                                                                  \
    do {                                                                                                               \
        if (lim == LIM_MC) stage_f32_batch_pair_kernel<F, LIM_MC><<<grid, block, 0, st>>>(cases, work, n_work, stage, parity); \
        else if (lim == LIM_MINMOD) stage_f32_batch_pair_kernel<F, LIM_MINMOD><<<grid, block, 0, st>>>(cases, work, n_work, stage, parity); \
        else stage_f32_batch_pair_kernel<F, LIM_NONE><<<grid, block, 0, st>>>(cases, work, n_work, stage, parity);      \
    } while (0)
    if (a_.flux == FLUX_RUSANOV) EFVM_BPAIR(FLUX_RUSANOV);
    else if (a_.flux == FLUX_HLLC) EFVM_BPAIR(FLUX_HLLC);
    else EFVM_BPAIR(FLUX_HLLC_TEXTBOOK);
#undef EFVM_BPAIR
    return hipGetLastError();
}

hipError_t launch_stage_f32_batch_pair23(const BatchCase* cases, const StageArgs& a_, const int2* work, int n_work,
                                         int parity, hipStream_t st) {
    const dim3 grid(n_work), block(2 * WAVE);
    const int lim = a_.muscl ? a_.limiter : LIM_NONE;
#define EFVM_B23_C(F, C)                                                                                               \
    do {                                                                                                               \
        if (lim == LIM_MC) stage_f32_batch_pair23_kernel<F, LIM_MC, C><<<grid, block, 0, st>>>(cases, work, parity);   \
        else if (lim == LIM_MINMOD) stage_f32_batch_pair23_kernel<F, LIM_MINMOD, C><<<grid, block, 0, st>>>(cases, work, parity); \
        else stage_f32_batch_pair23_kernel<F, LIM_NONE, C><<<grid, block, 0, st>>>(cases, work, parity);               \
    } while (0)
#define EFVM_B23(F) do { if (a_.fuse_cfl) EFVM_B23_C(F, true); else EFVM_B23_C(F, false); } while (0)
    if (a_.flux == FLUX_RUSANOV) EFVM_B23(FLUX_RUSANOV);
    else if (a_.flux == FLUX_HLLC) EFVM_B23(FLUX_HLLC);
    else EFVM_B23(FLUX_HLLC_TEXTBOOK);
#undef EFVM_B23
#undef EFVM_B23_C
    return hipGetLastError();
}

hipError_t launch_stage_f32_batch(const BatchCase* cases, const StageArgs& a_, const int2* work, int n_work, int stage,
                                  int parity, int blocks, int waves, hipStream_t st) {
    const dim3 grid(blocks), block(WAVE * waves);
#define EFVM_LAUNCH_B(F, L, X) stage_f32_batch_kernel<F, L, X><<<grid, block, 0, st>>>(cases, work, n_work, stage, parity)
    EFVM_F32_DISPATCH(EFVM_LAUNCH_B);
#undef EFVM_LAUNCH_B
    return hipGetLastError();
}

}  // namespace efvm
