// efvm_physics.h -- equation of state, numerical fluxes and limiters as inline
// host/device functions, templated on the arithmetic type.
//
// ref::* is the reference's exact operation order over five conserved
// components (rho, rho_u, rho_v, rho_w, rho_E), for the FP64 parity path.
// Every function names the reference lines it restates; compiled with
// -ffp-contract=off it reproduces the reference's rounding op for op.  The
// FP32 production schemes live in efvm_fast.h.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define EFVM_HD __host__ __device__ __forceinline__

namespace efvm {

enum FluxKind { FLUX_RUSANOV = 0, FLUX_HLLC = 1, FLUX_HLLC_TEXTBOOK = 2 };
enum LimiterKind { LIM_NONE = -1, LIM_MINMOD = 0, LIM_MC = 1 };
enum BcKind { BC_OUTFLOW = 0, BC_PERIODIC = 1, BC_SLIP = 2, BC_NOSLIP = 3, BC_SYMMETRY = 4,
              BC_SUP_INFLOW = 5, BC_SUB_INFLOW = 6, BC_PRESSURE_OUTLET = 7, BC_TABLE = 8 };

template <class T> struct C5 { T r, ru, rv, rw, re; };
template <class T> struct P6 { T r, u, v, w, p, t; };

namespace ref {

template <class T> EFVM_HD C5<T> add(C5<T> a, C5<T> b) { return {a.r + b.r, a.ru + b.ru, a.rv + b.rv, a.rw + b.rw, a.re + b.re}; }
template <class T> EFVM_HD C5<T> sub(C5<T> a, C5<T> b) { return {a.r - b.r, a.ru - b.ru, a.rv - b.rv, a.rw - b.rw, a.re - b.re}; }
template <class T> EFVM_HD C5<T> mul(C5<T> a, T s) { return {a.r * s, a.ru * s, a.rv * s, a.rw * s, a.re * s}; }

// ConsVar::velocity / kinetic_energy, crates/cfd_state/src/lib.rs:40-49
template <class T> EFVM_HD void velocity(const C5<T>& u, T v[3]) {
    T inv = T(1) / u.r;
    v[0] = u.ru * inv; v[1] = u.rv * inv; v[2] = u.rw * inv;
}
template <class T> EFVM_HD T kinetic(const C5<T>& u) { return T(0.5) * (u.ru * u.ru + u.rv * u.rv + u.rw * u.rw) / u.r; }
// IdealGas::pressure / sound_speed, crates/cfd_eos/src/lib.rs:61-68
template <class T> EFVM_HD T pressure(const C5<T>& u, T g) { return (g - T(1)) * (u.re - kinetic(u)); }
template <class T> EFVM_HD T sound(const C5<T>& u, T g) { T p = pressure(u, g); return sqrt(g * p / u.r); }
// IdealGas::cons_to_prim / prim_to_cons, crates/cfd_eos/src/lib.rs:44-59
template <class T> EFVM_HD P6<T> cons_to_prim(const C5<T>& u, T g, T R) {
    T inv = T(1) / u.r;
    T vu = u.ru * inv, vv = u.rv * inv, vw = u.rw * inv;
    T ke = T(0.5) * u.r * (vu * vu + vv * vv + vw * vw);
    T p = (g - T(1)) * (u.re - ke);
    T t = p * inv / R;
    return {u.r, vu, vv, vw, p, t};
}
template <class T> EFVM_HD C5<T> prim_to_cons(const P6<T>& p, T g) {
    T ke = T(0.5) * p.r * (p.u * p.u + p.v * p.v + p.w * p.w);
    T re = p.p / (g - T(1)) + ke;
    return {p.r, p.r * p.u, p.r * p.v, p.r * p.w, re};
}
// physical_flux_inline, crates/cfd_riemann/src/lib.rs:22-43
template <class T> EFVM_HD C5<T> phys_flux(const C5<T>& u, T g, int dir) {
    T p = pressure(u, g);
    T v[3]; velocity(u, v);
    T vn = v[dir];
    C5<T> f = {u.r * vn, u.ru * vn, u.rv * vn, u.rw * vn, (u.re + p) * vn};
    if (dir == 0) f.ru += p; else if (dir == 1) f.rv += p; else f.rw += p;
    return f;
}
// Rusanov::flux_inline, crates/cfd_riemann/src/lib.rs:222-240
template <class T> EFVM_HD C5<T> rusanov(const C5<T>& L, const C5<T>& R, T g, int dir) {
    C5<T> fl = phys_flux(L, g, dir), fr = phys_flux(R, g, dir);
    T vl[3], vr[3]; velocity(L, vl); velocity(R, vr);
    T al = sound(L, g), ar = sound(R, g);
    T lam = fmax(fabs(vl[dir]) + al, fabs(vr[dir]) + ar);
    return sub(mul(add(fl, fr), T(0.5)), mul(sub(R, L), T(0.5) * lam));
}
// star_state, crates/cfd_riemann/src/lib.rs:619-644
template <class T> EFVM_HD C5<T> star_state(const C5<T>& u, const T vel[3], T p, T un, T s, T ss, int dir) {
    T rs = u.r * (s - un) / (s - ss);
    T epm = u.re / u.r;
    T es = epm + (ss - un) * (ss + p / (u.r * (s - un)));
    C5<T> o;
    o.r = rs;
    o.ru = dir == 0 ? rs * ss : rs * vel[0];
    o.rv = dir == 1 ? rs * ss : rs * vel[1];
    o.rw = dir == 2 ? rs * ss : rs * vel[2];
    o.re = rs * es;
    return o;
}
// thornber_reconstruct, crates/cfd_riemann/src/lib.rs:456-506
template <class T> EFVM_HD void thornber(const C5<T>& L, const C5<T>& R, T g, C5<T>& Lc, C5<T>& Rc, T& al, T& ar) {
    T vl[3], vr[3]; velocity(L, vl); velocity(R, vr);
    T vlsq = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2];
    T vrsq = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
    T a_l = sound(L, g), a_r = sound(R, g);
    T vmax = fmax(vlsq, vrsq), amax = fmax(a_l, a_r);
    T z = (vmax >= amax * amax) ? T(1) : sqrt(vmax) / amax;
    T vlc[3], vrc[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        T avg = T(0.5) * (vl[d] + vr[d]);
        T dv = T(0.5) * z * (vl[d] - vr[d]);
        vlc[d] = avg + dv;
        vrc[d] = avg - dv;
    }
    T vlcsq = vlc[0] * vlc[0] + vlc[1] * vlc[1] + vlc[2] * vlc[2];
    T vrcsq = vrc[0] * vrc[0] + vrc[1] * vrc[1] + vrc[2] * vrc[2];
    Lc = {L.r, L.r * vlc[0], L.r * vlc[1], L.r * vlc[2], L.re + T(0.5) * L.r * (vlcsq - vlsq)};
    Rc = {R.r, R.r * vrc[0], R.r * vrc[1], R.r * vrc[2], R.re + T(0.5) * R.r * (vrcsq - vrsq)};
    al = a_l; ar = a_r;
}
// Hllc::flux_with_thornber, crates/cfd_riemann/src/lib.rs:395-440
template <class T> EFVM_HD C5<T> hllc(const C5<T>& L, const C5<T>& R, T g, int dir, bool use_thornber) {
    C5<T> lc, rc; T al, ar;
    if (use_thornber) thornber(L, R, g, lc, rc, al, ar);
    else { al = sound(L, g); ar = sound(R, g); lc = L; rc = R; }
    T vl[3], vr[3]; velocity(lc, vl); velocity(rc, vr);
    T pl = pressure(lc, g), pr = pressure(rc, g);
    T unl = vl[dir], unr = vr[dir];
    T sl = fmin(unl - al, unr - ar);
    T sr = fmax(unl + al, unr + ar);
    T fl = lc.r * (sl - unl);
    T fr = rc.r * (sr - unr);
    T ss = ((pr - pl) + fl * unl - fr * unr) / (fl - fr);
    if (sl >= T(0)) return phys_flux(lc, g, dir);
    if (ss >= T(0)) return add(phys_flux(lc, g, dir), mul(sub(star_state(lc, vl, pl, unl, sl, ss, dir), lc), sl));
    if (sr >= T(0)) return add(phys_flux(rc, g, dir), mul(sub(star_state(rc, vr, pr, unr, sr, ss, dir), rc), sr));
    return phys_flux(rc, g, dir);
}
template <class T> EFVM_HD C5<T> num_flux(int scheme, const C5<T>& L, const C5<T>& R, T g, int dir) {
    if (scheme == FLUX_RUSANOV) return rusanov(L, R, g, dir);
    return hllc(L, R, g, dir, scheme == FLUX_HLLC);
}
// minmod / monotonized_central, crates/cfd_solver/src/lib.rs:170-192
template <class T> EFVM_HD T minmod(T a, T b) {
    if (a * b <= T(0)) return T(0);
    return fabs(a) < fabs(b) ? a : b;
}
template <class T> EFVM_HD T mc(T a, T b) {
    if (a * b <= T(0)) return T(0);
    T s = a > T(0) ? T(1) : T(-1);
    return s * fmin(fmin(T(2) * fabs(a), T(2) * fabs(b)), T(0.5) * fabs(a + b));
}
template <class T> EFVM_HD T limit(int lim, T a, T b) { return lim == LIM_MC ? mc(a, b) : minmod(a, b); }
// Sutherland viscosity and Prandtl conductivity, crates/cfd_eos/src/lib.rs:75-87
template <class T> EFVM_HD T sutherland(T t) {
    const T MU_REF = T(1.827e-5), T_REF = T(291.15), S = T(120.0);
    return MU_REF * pow(t / T_REF, T(1.5)) * (T_REF + S) / (t + S);
}
template <class T> EFVM_HD T conductivity(T mu, T g, T R) { T cp = g * R / (g - T(1)); return mu * cp / T(0.72); }
}  // namespace ref
}  // namespace efvm
