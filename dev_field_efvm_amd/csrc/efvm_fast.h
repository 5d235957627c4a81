// efvm_fast.h -- FP32 production numerics for the 2-D path (rho_w == 0 dropped).
//
// Same schemes as the reference (HLLC with Thornber's low-Mach correction,
// crates/cfd_riemann/src/lib.rs:395-506; Rusanov :222-240; minmod / MC limiters,
// crates/cfd_solver/src/lib.rs:170-192), rewritten for the VALU:
//  * per face state one reciprocal (1/rho) and one reciprocal square root
//    (1/a, with a = (gamma p/rho) * rsq), shared by the flux that uses it;
//  * Thornber's z = min(|v|_max / a_max, 1) uses the stored 1/a;
//  * HLLC as F_K + S_K (U*_K - U_K) with U*_K - U_K written out as one factor
//    Phi/S_K = (S* - u_K)/(S_K - S*) times [rho, rho S_K, rho v_t, h + f_K S*]
//    (Toro sec. 10.4), S* and 1/(S_K - S*) from one reciprocal;
//  * upwind side chosen by selects; the star flux taken by select so an out-of-fan
//    inf/NaN never leaks into the flux.
// Mathematically identical to the reference; not bitwise.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace efvm {
namespace fast {

template <class T> struct C4 { T r, ru, rv, re; };

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#else
inline float rcp(float x) { return 1.0f / x; }
inline float rsq(float x) { return 1.0f / std::sqrt(x); }
inline float fsqrt(float x) { return std::sqrt(x); }
#endif

// minmod(a, b) is the median of (0, a, b); MC = sign(a) min(2|a|, 2|b|, |a+b|/2) when
// a b > 0, else 0, is the three-argument minmod(2a, 2b, (a+b)/2), i.e. minmod applied twice
// (on the device one med3 each; the host forms give the same values)
__host__ __device__ __forceinline__ float minmod(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(0.f, a, b);
#else
    float m = fabsf(a) < fabsf(b) ? a : b;
    return a * b > 0.f ? m : 0.f;
#endif
}
__host__ __device__ __forceinline__ float mc(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(0.f, 2.f * __builtin_amdgcn_fmed3f(0.f, a, b), 0.5f * (a + b));
#else
    float s = a + b;
    float m = fminf(2.f * fminf(fabsf(a), fabsf(b)), 0.5f * fabsf(s));
    return a * b > 0.f ? copysignf(m, s) : 0.f;
#endif
}
template <int LIM> __host__ __device__ __forceinline__ float limit(float a, float b) {
    if (LIM == 1) return mc(a, b);
    return minmod(a, b);
}

// Face state with the quantities every flux needs.
struct Prep {
    float r, ru, rv, re;   // conserved
    float u, v, p;         // velocity, pressure
    float q;               // |v|^2
    float a, ia, ir;       // sound speed, 1/a, 1/rho
};
__host__ __device__ __forceinline__ Prep prep(const C4<float>& q, float g, float gm1) {
    Prep o;
    o.r = q.r; o.ru = q.ru; o.rv = q.rv; o.re = q.re;
    o.ir = rcp(q.r);
    o.u = q.ru * o.ir;
    o.v = q.rv * o.ir;
    const float k = q.ru * o.u + q.rv * o.v;   // rho |v|^2
    o.q = k * o.ir;
    o.p = gm1 * (q.re - 0.5f * k);
    float x = g * o.p * o.ir;          // a^2
    o.ia = rsq(x);
    o.a = x * o.ia;
    return o;
}

// Face state from primitives (rho, u, v, p), as stored by the learned reconstruction.
__host__ __device__ __forceinline__ Prep prep_prim(float r, float u, float v, float p, float g, float igm1) {
    Prep o;
    o.r = r; o.u = u; o.v = v; o.p = p;
    o.ru = r * u;
    o.rv = r * v;
    o.q = u * u + v * v;
    o.re = p * igm1 + 0.5f * r * o.q;
    o.ir = rcp(r);
    float x = g * p * o.ir;
    o.ia = rsq(x);
    o.a = x * o.ia;
    return o;
}

// c ? a : b as a bit select.  Written with masks so that the seven selections on
// one condition stay lane selects: as conditional expressions they are merged
// into one divergent branch that copies every value on both paths.
__host__ __device__ __forceinline__ float pick(bool c, float a, float b) {
    const unsigned m = 0u - (unsigned)c;
    return __builtin_bit_cast(float, (__builtin_bit_cast(unsigned, a) & m) | (__builtin_bit_cast(unsigned, b) & ~m));
}

// HLLC (with Thornber's low-Mach velocity blend when THORNBER) through a face normal to axis
// DIR, from the upwind side K of the contact: F = F_K + Phi [rho_K, rho_K S_K, rho_K v_t,
// h_K + f_K S*], Phi = S_K (S* - u_K) / (S_K - S*) inside the wave fan (S_L < 0 <= S* or
// S* < 0 <= S_R), Phi = 0 outside it.  This is F*_K = F_K + S_K (U*_K - U_K) (Toro sec. 10.4)
// with U*_K - U_K = (S* - u_K)/(S_K - S*) [rho_K, rho_K S_K, rho_K v_t, h_K + f_K S*] written
// out, f_K = rho_K (S_K - u_K), h_K = E_K + p_K.  gg = gamma/(gamma-1): with Thornber's
// velocities the internal energy is kept, so h_K = gg p_K + rho_K |v'_K|^2 / 2.
template <int DIR, bool THORNBER>
__host__ __device__ __forceinline__ C4<float> hllc(const Prep& L, const Prep& R, float gg) {
    float ul = L.u, vl = L.v, ur = R.u, vr = R.v;
    if (THORNBER) {
        const float vm = fsqrt(fmaxf(L.q, R.q));
        // min(vm / a_L, vm / a_R): vm >= 0 and rounding is monotone, so scaling the
        // smaller 1/a gives the same value as the smaller of the two products
        const float z = fminf(vm * fminf(L.ia, R.ia), 1.f);
        // avg +/- z (v_L - v_R)/2  ==  v_L - w dv,  v_R + w dv  with w = (1 - z)/2, dv = v_L - v_R
        const float w = fmaf(-0.5f, z, 0.5f);
        const float du = ul - ur, dv = vl - vr;
        ul = fmaf(-w, du, ul); ur = fmaf(w, du, ur); vl = fmaf(-w, dv, vl); vr = fmaf(w, dv, vr);
    }
    const float unl = DIR == 0 ? ul : vl, unr = DIR == 0 ? ur : vr;
    const float sl = fminf(unl - L.a, unr - R.a);
    const float sr = fmaxf(unl + L.a, unr + R.a);
    const float fl = L.r * (sl - unl), fr = R.r * (sr - unr);
    // S* = N / Dn with Dn = fl - fr < 0 (fl < 0 < fr for positive rho, a), so the
    // contact side is known from the sign of N alone: S* >= 0  <=>  N <= 0.
    // One reciprocal then gives both S* and D = 1/(S_K - S*):
    //   X = S_K Dn - N = Dn (S_K - S*),  r = 1/(Dn X),  S* = N X r,  D = Dn^2 r.
    const float N = (R.p - L.p) + fl * unl - fr * unr;
    const float Dn = fl - fr;
    const bool left = N <= 0.f;
    const float rK = pick(left, L.r, R.r), uK = pick(left, ul, ur), vK = pick(left, vl, vr);
    const float pK = pick(left, L.p, R.p), sK = pick(left, sl, sr), fK = pick(left, fl, fr);
    const float unK = DIR == 0 ? uK : vK, utK = DIR == 0 ? vK : uK;
    const float hK = THORNBER ? fmaf(0.5f * rK, uK * uK + vK * vK, gg * pK) : pick(left, L.re, R.re) + pK;
    const float X = sK * Dn - N;
    const float rr = rcp(Dn * X);
    const float ss = N * X * rr;
    const float D = Dn * Dn * rr;
    // left: in the fan when S_L < 0; right: when S_R >= 0 -- one compare of S_K against the
    // side flag.  Outside the fan Phi and S* are taken as 0 by lane selects, so an out-of-fan
    // inf/NaN of the star terms never reaches the flux.
    const bool fan = (sK < 0.f) == left;
    const float phi = pick(fan, sK * (ss - unK) * D, 0.f);
    const float sf = pick(fan, ss, 0.f);
    C4<float> f;
    const float m = rK * (unK + phi);                  // rho_K u_K + Phi rho_K
    const float fn = m * unK + (pK + phi * fK);        // rho_K u_K^2 + p_K + Phi rho_K S_K
    f.r = m;
    f.ru = DIR == 0 ? fn : m * utK;
    f.rv = DIR == 1 ? fn : m * utK;
    f.re = hK * unK + phi * (hK + fK * sf);
    return f;
}

template <int DIR>
__host__ __device__ __forceinline__ C4<float> rusanov(const Prep& L, const Prep& R) {
    const float unl = DIR == 0 ? L.u : L.v, unr = DIR == 0 ? R.u : R.v;
    const float lam = fmaxf(fabsf(unl) + L.a, fabsf(unr) + R.a);
    C4<float> f;
    f.r = 0.5f * (L.r * unl + R.r * unr) - 0.5f * lam * (R.r - L.r);
    f.ru = 0.5f * (L.ru * unl + R.ru * unr + (DIR == 0 ? L.p + R.p : 0.f)) - 0.5f * lam * (R.ru - L.ru);
    f.rv = 0.5f * (L.rv * unl + R.rv * unr + (DIR == 1 ? L.p + R.p : 0.f)) - 0.5f * lam * (R.rv - L.rv);
    f.re = 0.5f * ((L.re + L.p) * unl + (R.re + R.p) * unr) - 0.5f * lam * (R.re - L.re);
    return f;
}

template <int FLUX, int DIR>
__host__ __device__ __forceinline__ C4<float> flux(const Prep& L, const Prep& R, float gg) {
    if (FLUX == 0) return rusanov<DIR>(L, R);
    return hllc<DIR, FLUX == 1>(L, R, gg);
}

}  // namespace fast
}  // namespace efvm
