// efvm_mlp.hip -- learned face reconstruction on gfx950.
//
//  * faces_kernel: for every face of a periodic grid, the two face sides' stencil
//    features, the MLP on bf16 16x16x32 matrix cores (FP32 accumulation), the
//    FP32 skip and offsets, the per-face positivity / TVD check with fallback to
//    the solver's MUSCL (conserved variables, MC limiter), and the final face
//    primitives for the learned stage kernel.
//  * faces_light_kernel / faces_list_kernel: the bypass mode as separate passes (open-gate
//    faces and their fluxes, then the deferred faces through the network).
//  * stage_light_kernel / pending_kernel: the bypass mode as one fused stage (faces, fluxes,
//    update and CFL in LDS-linked row tiles), then the cells next to a deferred face.
//  * train_kernel / reduce_kernel / adam_kernel: FP32 minibatch loss, gradient
//    (per-block partials reduced in a fixed order, so gradients are
//    deterministic) and the Adam update.
//  * coarse / face-target / sample kernels: training samples from a fine solver
//    state, box-filtered on the device (the data never leaves HBM).
//  * infer_kernel: the matrix-core MLP on given features (for checking).
//
// The reference has no learned part; the layout and formulas are in efvm_mlp.h.
#include "efvm_stage.h"
#include "efvm_fast.h"
#include "efvm_mlp.h"
#include "efvm_cfl_f32.h"

namespace efvm {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pack2(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// tanh-form GELU: x * sigmoid(2 c (x + a x^3)) = x / (1 + 2^(x (K0 + K1 x^2)))
__device__ __forceinline__ float gelu_f(float x) {
    constexpr float K0 = -2.f * GELU_C * 1.4426950408889634f;
    constexpr float K1 = K0 * GELU_A;
    const float e = __builtin_amdgcn_exp2f(x * fmaf(K1, x * x, K0));
    return x * fast::rcp(1.f + e);
}

// Per-lane matrix-core operands of the three layers (see Packed).
struct Regs {
    bf16x8 a1[4], a2[4][2], a3[2];
    f32x4 c2[4];                 // layer-2 bias (layer 1's rides on the constant feature 20)
};

__device__ __forceinline__ void load_regs(const Packed* __restrict__ P, int lane, Regs& R) {
    const int g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        R.a1[t] = __builtin_bit_cast(bf16x8, *(const u32x4*)P->w1[t][lane]);
#pragma unroll
        for (int s = 0; s < 2; s++) R.a2[t][s] = __builtin_bit_cast(bf16x8, *(const u32x4*)P->w2[t][s][lane]);
        R.c2[t] = f32x4{P->b2[16 * t + 4 * g], P->b2[16 * t + 4 * g + 1], P->b2[16 * t + 4 * g + 2], P->b2[16 * t + 4 * g + 3]};
    }
#pragma unroll
    for (int s = 0; s < 2; s++) R.a3[s] = __builtin_bit_cast(bf16x8, *(const u32x4*)P->w3[s][lane]);
}

// gelu of two accumulator tiles -> the bf16 B operand of the next layer
__device__ __forceinline__ bf16x8 act_pack(const f32x4& lo, const f32x4& hi) {
    const u32x4 v = {pack2(gelu_f(lo[0]), gelu_f(lo[1])), pack2(gelu_f(lo[2]), gelu_f(lo[3])),
                     pack2(gelu_f(hi[0]), gelu_f(hi[1])), pack2(gelu_f(hi[2]), gelu_f(hi[3]))};
    return __builtin_bit_cast(bf16x8, v);
}

// MLP of 16 face sides: bx = the lane's 8 input features (layer-1 B operand);
// returns the layer-3 accumulator (without b3): element r of lane (g, j) = output r of side j.
__device__ __forceinline__ f32x4 mlp16(const Regs& R, bf16x8 bx) {
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 d1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) d1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a1[t], bx, zero, 0, 0, 0);
    const bf16x8 h0 = act_pack(d1[0], d1[1]), h1 = act_pack(d1[2], d1[3]);
    f32x4 d2[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        d2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a2[t][0], h0, R.c2[t], 0, 0, 0);
        d2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a2[t][1], h1, d2[t], 0, 0, 0);
    }
    const bf16x8 k0 = act_pack(d2[0], d2[1]), k1 = act_pack(d2[2], d2[3]);
    f32x4 d3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a3[0], k0, zero, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a3[1], k1, d3, 0, 0, 0);
}

constexpr int SLOT = 20;   // dwords per face-side slot in LDS (16 used; 20 keeps b128 reads conflict-free)
// FP32 skip weights staged in LDS (read as broadcasts, so they occupy no registers)
constexpr int SK_L = 0, SK_BL = NOUT * NIN, SK_B3 = SK_BL + NOUT, SK_N = SK_B3 + NOUT;

__device__ __forceinline__ void stage_skip(const Packed* __restrict__ P, float* skip) {
    for (int i = threadIdx.x; i < SK_N; i += blockDim.x)
        skip[i] = i < SK_BL ? P->L[i / NIN][i % NIN] : (i < SK_B3 ? P->bL[i - SK_BL] : P->b3[i - SK_B3]);
    __syncthreads();
}

// the gate of one output: sigmoid(z) = 1 / (1 + 2^(-z log2 e))
__device__ __forceinline__ float gate(float z) {
    return fast::rcp(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

// Features of one oriented stencil ws[k][v] (k = 0..4 <-> cells -2..2).  Writes the 20
// features as bf16 into the slot (dwords 0..9, plus the constant feature 20 = 1 in dword
// 10), returns sigma, the centre, the bound tolerance, the FP32 central-slope base yb and
// the FP32 learned correction sk = L x + bL (which the MLP's gate scales).
template <bool SLOT_OUT = true>
__device__ __forceinline__ void featurize(const float ws[5][4], float gamma, const float* skip, bool own4, uint32_t* slot,
                                          float sig[4], float cen[4], float yb[4], float sk[4], float tol[4]) {
    float x[20];
    const float a0 = fast::fsqrt(gamma * ws[2][3] * fast::rcp(ws[2][0]));
    const float scale[4] = {ws[2][0], a0, a0, ws[2][3]};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float c = ws[2][v];
        float m = FLOOR * scale[v];
#pragma unroll
        for (int k = 0; k < 5; k++)
            if (k != 2) m = fmaxf(m, fabsf(ws[k][v] - c));
        const float is = fast::rcp(m);
#pragma unroll
        for (int k = 0; k < 5; k++)
            if (k != 2) x[4 * k + v] = (ws[k][v] - c) * is;
        x[8 + v] = m * fast::rcp(m + X0C * scale[v]);
        sig[v] = m;
        cen[v] = c;
        tol[v] = FLOOR * scale[v];
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        float s = skip[SK_BL + v];
        if (own4) {
            // the nonzero terms of the dense sum below, in its order: the same result bit for bit
#pragma unroll
            for (int k = 0; k < 5; k++)
                if (k != 2) s = fmaf(skip[SK_L + v * NIN + 4 * k + v], x[4 * k + v], s);
        } else {
#pragma unroll
            for (int f = 0; f < 20; f++) s = fmaf(skip[SK_L + v * NIN + f], x[f], s);
        }
        yb[v] = 0.25f * (x[12 + v] - x[4 + v]);
        sk[v] = s;
    }
    if (!SLOT_OUT) return;
    const u32x4 q0 = {pack2(x[0], x[1]), pack2(x[2], x[3]), pack2(x[4], x[5]), pack2(x[6], x[7])};
    const u32x4 q1 = {pack2(x[8], x[9]), pack2(x[10], x[11]), pack2(x[12], x[13]), pack2(x[14], x[15])};
    *(u32x4*)(slot) = q0;
    *(u32x4*)(slot + 4) = q1;
    slot[8] = pack2(x[16], x[17]);
    slot[9] = pack2(x[18], x[19]);
    slot[10] = pack2(1.f, 0.f);      // constant feature 20 carries the layer-1 bias
}

// Positivity and the TVD bound relaxed at smooth extrema: every variable of the oriented
// face value within [min(w0, w1) - d, max(w0, w1) + d], d = relax |minmod(curvatures)| + tol
// (relax = 1/2; tol = the feature floor, below which offsets are not resolved).
__device__ __forceinline__ bool side_ok(const float ws[5][4], const float f[4], const float tol[4],
                                        float relax = 0.5f) {
    bool ok = f[0] > 0.f && f[3] > 0.f;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float w0 = ws[2][v], w1 = ws[3][v];
        const float c0 = w1 - 2.f * w0 + ws[1][v];
        const float c1 = ws[4][v] - 2.f * w1 + w0;
        const float d = (c0 * c1 > 0.f ? relax * fminf(fabsf(c0), fabsf(c1)) : 0.f) + tol[v];
        ok = ok && f[v] >= fminf(w0, w1) - d && f[v] <= fmaxf(w0, w1) + d;
    }
    return ok;
}

// The open-gate face value of one oriented side with the 4-tap correction: fo = w_c +
// sum_k t_k (w_k - w_c) over the stencil cells k = 0, 1, 3, 4 (taps tp[0..3][v]: L's own-variable
// weights, the central slope's -1/4 and +1/4 folded into the inner two).  With the gate open and
// no bias the feature normalisation cancels, so this is the face value the network path
// computes for phi = 1, up to rounding.
__device__ __forceinline__ void light_value(const float ws[5][4], const float (&tp)[4][4], float fo[4]) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float c = ws[2][v];
        float d = tp[0][v] * (ws[0][v] - c);
        d = fmaf(tp[1][v], ws[1][v] - c, d);
        d = fmaf(tp[2][v], ws[3][v] - c, d);
        d = fmaf(tp[3][v], ws[4][v] - c, d);
        fo[v] = c + d;
    }
}

// second difference of a variable at a cell, one symmetric form for every face it bounds
__device__ __forceinline__ float curvature(float wm, float w0, float wp) { return (wm + wp) - 2.f * w0; }

// The strict bound (half the check's relaxation) for both sides of one face, in the frame whose
// normal points from the left cell a to the right cell b (variables rho, u_n, u_t, p): ca, cb =
// the cells' curvatures along the normal, fl / fr = the two face values in that frame, tl / tr =
// the tolerances of the sides' own cells (FLOOR x (rho, a0, a0, p)).  A side passes when its
// density and pressure are positive and every variable lies within [min(a, b) - d, max(a, b) + d],
// d = relax |minmod(ca, cb)| + tol: side_ok's bound, written once per face (the right side's
// oriented test negates its normal velocity and the bounds alike, so the two are the same test).
__device__ __forceinline__ void face_test(const float a[4], const float b[4], const float ca[4], const float cb[4],
                                          const float fl[4], const float fr[4], const float tl[4], const float tr[4],
                                          bool& okl, bool& okr) {
    okl = fl[0] > 0.f && fl[3] > 0.f;
    okr = fr[0] > 0.f && fr[3] > 0.f;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float lo = fminf(a[v], b[v]), hi = fmaxf(a[v], b[v]);
        const float m = ca[v] * cb[v] > 0.f ? 0.25f * fminf(fabsf(ca[v]), fabsf(cb[v])) : 0.f;
        const float dl = m + tl[v], dr = m + tr[v];
        okl = okl && fl[v] >= lo - dl && fl[v] <= hi + dl;
        okr = okr && fr[v] >= lo - dr && fr[v] <= hi + dr;
    }
}
__device__ __forceinline__ void tol_of(float rho, float a0, float p, float t[4]) {
    t[0] = FLOOR * rho;
    t[1] = FLOOR * a0;
    t[2] = FLOOR * a0;
    t[3] = FLOOR * p;
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// The Riemann flux through one face from the lab-frame primitives (rho, u, v, p) of its
// left / lower (L) and right / upper (R) sides.
template <int FLUX, int DIR>
__device__ __forceinline__ fast::C4<float> face_flux_t(float gm, float igm1, float gg, const float L[4], const float R[4]) {
    const fast::Prep pl = fast::prep_prim(L[0], L[1], L[2], L[3], gm, igm1);
    const fast::Prep pr = fast::prep_prim(R[0], R[1], R[2], R[3], gm, igm1);
    return fast::flux<FLUX, DIR>(pl, pr, gg);
}
__device__ __forceinline__ fast::C4<float> face_flux(const FaceArgs& a, int axis, const float L[4], const float R[4]) {
    const float gm = a.gamma, igm1 = 1.f / (a.gamma - 1.f), gg = gm * igm1;
    if (axis == 0)
        return a.flux == 0 ? face_flux_t<0, 0>(gm, igm1, gg, L, R)
                           : (a.flux == 1 ? face_flux_t<1, 0>(gm, igm1, gg, L, R) : face_flux_t<2, 0>(gm, igm1, gg, L, R));
    return a.flux == 0 ? face_flux_t<0, 1>(gm, igm1, gg, L, R)
                       : (a.flux == 1 ? face_flux_t<1, 1>(gm, igm1, gg, L, R) : face_flux_t<2, 1>(gm, igm1, gg, L, R));
}

// Write one face: its primitives (out_mode 0) or its Riemann flux (out_mode 1).  L, R are
// the lab-frame primitives (rho, u, v, p) of the left / lower and right / upper sides.
__device__ __forceinline__ void put_face(const FaceArgs& a, int axis, long cell, long np, const float L[4], const float R[4]) {
    if (a.out_mode == 1) {
        const fast::C4<float> f = face_flux(a, axis, L, R);
        float* Fo = a.F + (long)(axis * 4) * np + cell;
        Fo[0] = f.r; Fo[np] = f.ru; Fo[2 * np] = f.rv; Fo[3 * np] = f.re;
    } else {
        float* Fl = a.F + (long)(axis * 2) * 4 * np + cell;
        float* Fr = Fl + 4 * np;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            Fl[v * np] = L[v];
            Fr[v * np] = R[v];
        }
    }
}

// Intra-wave LDS hand-off: the slots a wave writes are read only by its own lanes, and a
// wave's LDS operations execute in order, so only compiler reordering must be fenced.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The 6 cells m = -2..3 straddling the lane's face along `axis`: conserved q, primitives w.
__device__ __forceinline__ void load_raw(const FaceArgs& a, int axis, int row, int cc, float q[6][4]) {
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const long idx = axis == 0 ? (long)row * a.nx + wrap(cc + m - 2, a.nx) : (long)wrap(row + m - 2, a.ny) * a.nx + cc;
#pragma unroll
        for (int v = 0; v < 4; v++) q[m][v] = a.U[v][idx];
    }
}
// primitives (rho, u, v, p) of one cell from its conserved variables; ir = 1 / rho
__device__ __forceinline__ void prim1(const float q[4], float gm1, float w[4], float& ir) {
    ir = fast::rcp(q[0]);
    w[0] = q[0];
    w[1] = q[1] * ir;
    w[2] = q[2] * ir;
    w[3] = gm1 * (q[3] - 0.5f * (q[1] * w[1] + q[2] * w[2]));
}
__device__ __forceinline__ void to_prim(const float q[6][4], float gm1, float w[6][4]) {
#pragma unroll
    for (int m = 0; m < 6; m++) {
        float ir;
        prim1(q[m], gm1, w[m], ir);
    }
}
__device__ __forceinline__ void load_cells(const FaceArgs& a, int axis, int row, int cc, float q[6][4], float w[6][4]) {
    load_raw(a, axis, row, cc, q);
    to_prim(q, a.gamma - 1.f, w);
}

// task -> (axis, row, column of the lane's face); valid = the column exists
__device__ __forceinline__ void task_face(const FaceArgs& a, int task, int lane, int& axis, int& row, int& cc, bool& valid) {
    const int half = a.ny * a.n_chunks;
    axis = task >= half ? 1 : 0;
    const int rem = task - axis * half;
    row = rem / a.n_chunks;
    const int ci = (rem - row * a.n_chunks) * 64 + lane;
    valid = ci < a.nx;
    cc = valid ? ci : a.nx - 1;
}

// oriented stencil of side sd: 0 = the cell m=0 toward +, 1 = the cell m=1 toward -
__device__ __forceinline__ void orient(const float w[6][4], int axis, int sd, float ws[5][4]) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int m = sd == 0 ? k : 5 - k;   // k = 0..4 <-> cells -2..2 of the side's owner
        const float un = axis == 0 ? w[m][1] : w[m][2];
        const float ut = axis == 0 ? w[m][2] : w[m][1];
        ws[k][0] = w[m][0];
        ws[k][1] = sd == 0 ? un : -un;
        ws[k][2] = ut;
        ws[k][3] = w[m][3];
    }
}

// Face kernels.  A task is 64 faces: in PASS_ALL / PASS_LIGHT 64 consecutive faces of
// one row (x faces, east of cells chunk*64 + lane) or of one face row (y faces, north
// of those cells); in PASS_LIST 64 entries of the deferred-face list.  Each lane loads
// the 6 cells straddling its face and builds both sides' features (left side = its
// cell oriented +, right side = the neighbour oriented -).
//  * PASS_ALL: the network gates every face side (with a.bypass, only the sides whose
//    ungated value fails the TVD bound with half its relaxation), in 8 matrix-core
//    blocks of 16 sides; then the checks and the MUSCL fallback.  Register-bound by the
//    network's operands: 2 waves per SIMD.
//  * PASS_LIGHT (bypass stage, no network): faces whose two sides pass the strict bound
//    are final with the gate fully open (the strict bound implies the check); faces
//    with a side that fails it are appended to the deferred list.  No network
//    registers, so it runs at full occupancy.
//  * PASS_LIST: the deferred faces, as PASS_ALL with bypass; a persistent grid walks the
//    list up to the count PASS_LIGHT left in device memory.
// All passes write the face primitives (16 planes) the stage kernel reads.
enum { PASS_ALL = 0, PASS_LIST = 2 };

// One face per lane (axis, row, column of its west / south cell; `valid` lanes only count)
// through the network: both sides' features, the gate (with `bypass`, only for the sides
// whose ungated value fails the TVD bound with half its relaxation), the checks and the
// MUSCL fallback.  Returns the lab-frame primitives of the left / lower (L) and right /
// upper (Rr) face states, whether the face fell back, and which sides the network gated.
// Wave-uniform control: every lane of the wave calls it together.
__device__ __forceinline__ void net_face(const FaceArgs& a, const float* skip, uint32_t (*feat)[SLOT], const Regs& R,
                                         bool bypass, bool own4, int axis, int row, int cc, bool valid, float L[4],
                                         float Rr[4], bool& fell, bool& need0, bool& need1) {
    const int lane = lane_id();
    const int g = lane >> 4, jl = lane & 15;
    const float gm = a.gamma, gm1 = a.gamma - 1.f;
    float sig[2][4], cen[2][4], y[2][4], sk[2][4], tol[2][4];
    need0 = true;
    need1 = true;   // the side's gate is evaluated by the network
    {
        float q[6][4], w[6][4], ws[5][4], f1[4];
        load_cells(a, axis, row, cc, q, w);
        orient(w, axis, 0, ws);
        featurize(ws, gm, skip, own4, feat[lane], sig[0], cen[0], y[0], sk[0], tol[0]);
        if (bypass) {
#pragma unroll
            for (int v = 0; v < 4; v++) f1[v] = fmaf(sig[0][v], y[0][v] + sk[0][v], cen[0][v]);
            need0 = !side_ok(ws, f1, tol[0], 0.25f);
        }
        orient(w, axis, 1, ws);
        featurize(ws, gm, skip, own4, feat[64 + lane], sig[1], cen[1], y[1], sk[1], tol[1]);
        if (bypass) {
#pragma unroll
            for (int v = 0; v < 4; v++) f1[v] = fmaf(sig[1][v], y[1][v] + sk[1][v], cen[1][v]);
            need1 = !side_ok(ws, f1, tol[1], 0.25f);
        }
    }
    float fo[2][4];
    bool ok = true;
    float q[6][4], w[6][4];
    {
        wave_sync();
        const unsigned long long nm0 = __ballot(valid && need0), nm1 = __ballot(valid && need1);
        f32x4 o0 = {0.f, 0.f, 0.f, 0.f}, o1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int b = 0; b < 8; b++) {
            const int sd = b >> 2, qd = b & 3;
            if ((((sd == 0 ? nm0 : nm1) >> (16 * qd)) & 0xffffull) == 0ull) continue;   // no side needs the gate
            const bf16x8 bx = __builtin_bit_cast(bf16x8, *(const u32x4*)&feat[sd * 64 + qd * 16 + jl][4 * g]);
            const f32x4 d3 = mlp16(R, bx);
            if (g == qd) {
                if (sd == 0) o0 = d3; else o1 = d3;
            }
        }
        wave_sync();
        // face values (oriented), checks
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float b3 = skip[SK_B3 + v];
            const float p0 = need0 ? gate(o0[v] + b3) : 1.f, p1 = need1 ? gate(o1[v] + b3) : 1.f;
            fo[0][v] = fmaf(sig[0][v], fmaf(p0, sk[0][v], y[0][v]), cen[0][v]);
            fo[1][v] = fmaf(sig[1][v], fmaf(p1, sk[1][v], y[1][v]), cen[1][v]);
        }
        load_cells(a, axis, row, cc, q, w);
        float ws[5][4];
        orient(w, axis, 0, ws);
        ok = side_ok(ws, fo[0], tol[0]);
        orient(w, axis, 1, ws);
        ok = side_ok(ws, fo[1], tol[1]) && ok;
    }
    {
        const float unl = fo[0][1], unr = -fo[1][1];
        L[0] = fo[0][0]; L[3] = fo[0][3];
        Rr[0] = fo[1][0]; Rr[3] = fo[1][3];
        L[1] = axis == 0 ? unl : fo[0][2];
        L[2] = axis == 0 ? fo[0][2] : unl;
        Rr[1] = axis == 0 ? unr : fo[1][2];
        Rr[2] = axis == 0 ? fo[1][2] : unr;
    }
    if (!ok) {
        // the solver's MUSCL: MC-limited conserved slopes, cells m = -1..2 (indices 1..4)
        float ql[4], qr[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float sl = fast::mc(q[3][v] - q[2][v], q[2][v] - q[1][v]);
            const float sr = fast::mc(q[4][v] - q[3][v], q[3][v] - q[2][v]);
            ql[v] = fmaf(0.5f, sl, q[2][v]);
            qr[v] = fmaf(-0.5f, sr, q[3][v]);
        }
        const float il = fast::rcp(ql[0]), irr = fast::rcp(qr[0]);
        L[0] = ql[0]; L[1] = ql[1] * il; L[2] = ql[2] * il;
        L[3] = gm1 * (ql[3] - 0.5f * (ql[1] * L[1] + ql[2] * L[2]));
        Rr[0] = qr[0]; Rr[1] = qr[1] * irr; Rr[2] = qr[2] * irr;
        Rr[3] = gm1 * (qr[3] - 0.5f * (qr[1] * Rr[1] + qr[2] * Rr[2]));
    }
    fell = !ok;
}

template <int PASS>   // PASS_ALL or PASS_LIST (PASS_LIGHT is faces_light_kernel)
__device__ __forceinline__ void faces_body(const FaceArgs& a, const float* skip, uint32_t (*feat)[SLOT], const Regs& R) {
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const long np = (long)a.nx * a.ny;
    const bool bypass = PASS != PASS_ALL || a.bypass;
    const bool own4 = a.w->own4 != 0;
    unsigned long long nfb = 0, ntot = 0, nnet = 0;
    const int half = a.ny * a.n_chunks;
    const int n_tasks = PASS == PASS_LIST ? (int)((*a.n_list + 63u) / 64u) : a.n_tasks;
    for (int task = blockIdx.x * 4 + wv; task < n_tasks; task += gridDim.x * 4) {
        int axis, row, cc;
        bool valid;
        if (PASS == PASS_LIST) {
            const unsigned e = (unsigned)task * 64u + (unsigned)lane;
            valid = e < *a.n_list;
            const unsigned f = valid ? a.list[e] : 0u;          // axis * nx*ny + row * nx + column
            axis = f >= (unsigned)np ? 1 : 0;
            const unsigned r = f - (unsigned)(axis * np);
            row = (int)(r / (unsigned)a.nx);
            cc = (int)(r - (unsigned)row * (unsigned)a.nx);
        } else {
            axis = task >= half ? 1 : 0;
            const int rem = task - axis * half;
            row = rem / a.n_chunks;
            const int ci = (rem - row * a.n_chunks) * 64 + lane;
            valid = ci < a.nx;
            cc = valid ? ci : a.nx - 1;
        }
        float L[4], Rr[4];   // lab-frame primitives of the left / right face states
        bool fell, need0, need1;
        net_face(a, skip, feat, R, bypass, own4, axis, row, cc, valid, L, Rr, fell, need0, need1);
        nnet += __popcll(__ballot(valid && need0)) + __popcll(__ballot(valid && need1));
        nfb += __popcll(__ballot(valid && fell));
        ntot += __popcll(__ballot(valid));
        if (valid) put_face(a, axis, (long)row * a.nx + cc, np, L, Rr);
    }
    if (a.stats && lane == 0 && (ntot || nnet)) {
        // the wave's own slot (kernels on one stream run one after another)
        unsigned long long* st = a.stats + 3 * (blockIdx.x * 4 + wv);
        st[0] += nfb;
        st[1] += ntot;
        st[2] += nnet;
    }
}

__device__ __forceinline__ void init_feat_slots(uint32_t (*feat)[SLOT]) {
    for (int e = lane_id(); e < 128; e += 64)
        for (int d = 10; d < SLOT; d++) feat[e][d] = 0u;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void faces_kernel(FaceArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t lds[4][128][SLOT];
    __shared__ float skip[SK_N];
    stage_skip(a.w, skip);
    init_feat_slots(lds[threadIdx.x >> 6]);
    Regs R;
    load_regs(a.w, lane_id(), R);
    faces_body<PASS_ALL>(a, skip, lds[threadIdx.x >> 6], R);
}

// PASS_LIGHT on its own: one side at a time, so only the stencil and the two face
// values stay live (no network operands, no per-side feature state).  OWN4 (the
// correction is a 4-tap stencil per variable, the shipped network's form) is a template
// parameter: the general featurisation's registers would otherwise set the occupancy.
template <bool OWN4>
__global__ __launch_bounds__(256) void faces_light_kernel(FaceArgs a) {
    __shared__ float skip[SK_N];
    stage_skip(a.w, skip);
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const float gm = a.gamma, gm1 = a.gamma - 1.f;
    const long np = (long)a.nx * a.ny;
    constexpr bool own4 = OWN4;
    const float (&tp)[4][4] = a.taps;
    unsigned long long ntot = 0;
    for (int task = blockIdx.x * 4 + wv; task < a.n_tasks; task += gridDim.x * 4) {
        int axis, row, cc;
        bool valid;
        task_face(a, task, lane, axis, row, cc, valid);
        float q[6][4], w[6][4], fo[2][4];
        load_cells(a, axis, row, cc, q, w);
        bool defer = false;
        if (own4) {
            // gate open, 4-tap correction, no bias: the normalisation cancels and the face value
            // is the linear stencil w0 + sum_k t_k (w_k - w0) with t = L + central slope
            float ws[5][4];
            orient(w, axis, 0, ws);
            light_value(ws, tp, fo[0]);
            float a[4], b[4], ca[4], cb[4], tl[4], tr[4], frp[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                a[v] = ws[2][v];
                b[v] = ws[3][v];
                ca[v] = curvature(ws[1][v], ws[2][v], ws[3][v]);
                cb[v] = curvature(ws[2][v], ws[3][v], ws[4][v]);
            }
            tol_of(ws[2][0], fast::fsqrt(gm * ws[2][3] * fast::rcp(ws[2][0])), ws[2][3], tl);
            tol_of(ws[3][0], fast::fsqrt(gm * ws[3][3] * fast::rcp(ws[3][0])), ws[3][3], tr);
            orient(w, axis, 1, ws);
            light_value(ws, tp, fo[1]);
            frp[0] = fo[1][0]; frp[1] = -fo[1][1]; frp[2] = fo[1][2]; frp[3] = fo[1][3];
            bool okl, okr;
            face_test(a, b, ca, cb, fo[0], frp, tl, tr, okl, okr);
            defer = !(okl && okr);
        } else {
#pragma unroll
            for (int sd = 0; sd < 2; sd++) {
                float ws[5][4], sig[4], cen[4], y[4], sk[4], tol[4];
                orient(w, axis, sd, ws);
                featurize<false>(ws, gm, skip, own4, nullptr, sig, cen, y, sk, tol);
#pragma unroll
                for (int v = 0; v < 4; v++) fo[sd][v] = fmaf(sig[v], y[v] + sk[v], cen[v]);   // gate fully open
                defer = defer || !side_ok(ws, fo[sd], tol, 0.25f);
            }
        }
        // defer the faces with a side near the bound (wave-aggregated append)
        defer = defer && valid;
        const unsigned long long dm = __ballot(defer);
        const long cell = (long)row * a.nx + cc;
        if (dm) {
            unsigned base = 0u;
            if (lane == 0) base = atomicAdd(a.n_list, (unsigned)__popcll(dm));
            base = __shfl(base, 0, WAVE);
            if (defer) a.list[base + (unsigned)__popcll(dm & ((1ull << lane) - 1ull))] = (unsigned)(axis * np + cell);
        }
        valid = valid && !defer;
        ntot += __popcll(__ballot(valid));
        if (valid) {
            // both sides pass the strict bound, hence the check: the face is final
            const float unl = fo[0][1], unr = -fo[1][1];
            const float L[4] = {fo[0][0], axis == 0 ? unl : fo[0][2], axis == 0 ? fo[0][2] : unl, fo[0][3]};
            const float Rr[4] = {fo[1][0], axis == 0 ? unr : fo[1][2], axis == 0 ? fo[1][2] : unr, fo[1][3]};
            put_face(a, axis, cell, np, L, Rr);
        }
    }
    if (a.stats && lane == 0 && ntot) a.stats[3 * (blockIdx.x * 4 + wv) + 1] += ntot;   // the wave's own slot
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void faces_list_kernel(FaceArgs a) {
    if (*a.n_list == 0u) return;    // nothing deferred (the usual case on smooth flow)
    __shared__ __attribute__((aligned(16))) uint32_t lds[4][128][SLOT];
    __shared__ float skip[SK_N];
    stage_skip(a.w, skip);
    init_feat_slots(lds[threadIdx.x >> 6]);
    Regs R;
    load_regs(a.w, lane_id(), R);
    faces_body<PASS_LIST>(a, skip, lds[threadIdx.x >> 6], R);
}

// ---------------------------------------------------------------------------
// Fused bypass stage (periodic grids, 4-tap correction).  One wave per strip of LIGHT_OUT
// columns and strip_h rows computes every face of its cells with the open gate, tests both
// sides against the strict bound, takes the Riemann flux and updates the cell in place of
// the light pass, the deferred-face pass and the flux-plane stage: the fluxes never leave
// the registers.  A face with a side that fails the strict bound is deferred as in the
// light pass; the cells on both sides of it are then left to pending_kernel, which gets
// the cell and a mask of its deferred faces through the list, and the fluxes of the
// cell's other faces through the flux planes.  Every value that reaches the state is
// computed by the expressions of the separate passes (light_value, face_test, face_flux_t, the update
// with its multiply-adds written out), so the two paths agree bit for bit.
// ---------------------------------------------------------------------------
static_assert(LIGHT_OUT == WAVE - 6, "x faces need the cells -2..+3 of the wave");

__device__ __forceinline__ float lane_from_left(float x) {    // lane i gets lane i-1's x (lane 0: 0)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_from_right(float x) {   // lane i gets lane i+1's x (lane 63: 0)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ int lane_from_left_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int lane_from_right_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true); }


// RK coefficients and dt of a stage (the flux-plane stage's)
struct StageCoef { double dtd; float dtix, dtiy, ca, cb; bool fuse; };
__device__ __forceinline__ StageCoef stage_coef(const StageArgs& a) {
    StageCoef k;
    k.dtd = a.fixed_dt > 0.0 ? a.fixed_dt : a.cfl / a.clock->lam[a.parity];
    const float dt = (float)k.dtd;
    k.fuse = a.fuse_cfl && (a.euler || a.stage == 2);
    if (a.euler || a.stage == 0) { k.ca = 0.f; k.cb = 1.f; }
    else if (a.stage == 1) { k.ca = 0.75f; k.cb = 0.25f; }
    else { k.ca = 1.f / 3.f; k.cb = 2.f / 3.f; }
    k.dtix = dt * (float)a.inv_x;
    k.dtiy = dt * (float)a.inv_y;
    return k;
}
// new state of one variable from its four face fluxes (west, east, south, north), the
// stage input u and U^n u0; the flux-plane stage's expression
__device__ __forceinline__ float cell_update(const StageCoef& k, float fw, float fe, float gs, float gn, float u, float u0) {
    const float d = fmaf(fw - fe, k.dtix, (gs - gn) * k.dtiy);
    return k.ca != 0.f ? fmaf(k.cb, (u - u0) + d, u0) : u + d;
}

// One workgroup of LIGHT_ROWS + 1 waves per tile of LIGHT_OUT columns by LIGHT_ROWS rows j0..
// (grid-stride over tiles).  Phase A: wave w builds the face (j0-1+w, j0+w) whole (the north side
// of its lower row, the south side of its upper row, the Riemann flux) and puts the flux in LDS;
// waves w < LIGHT_ROWS also build the x faces of row j0+w.  Phase B: wave w updates row j0+w from
// the LDS fluxes of faces w and w+1.  No state runs from row to row and no side is built twice.
// A stage is latency- and issue-bound at the grid sizes the learned runs use: a 512^2 stage is
// 1539 workgroups of 4 short waves (79 VGPRs, 6 waves per SIMD), all resident at once; strips
// marching rows per wave (2 waves per SIMD, one long dependent chain each) measured slower than
// the separate passes, tiles of 4 or 5 rows took a second round of workgroups.  (Folding the CFL
// maxima in the last workgroup to finish, through a counter every workgroup increments, measured
// 30 us slower per stage than the separate reduction launch: same-address atomics from
// workgroups that finish together serialise.)
template <int FLUX>
__global__ __launch_bounds__(64 * (LIGHT_ROWS + 1)) void stage_light_kernel(StageArgs a, FaceArgs fa) {
    constexpr int NW = LIGHT_ROWS + 1;   // waves per workgroup
    __shared__ float4 sG[NW][WAVE];      // flux through face (j0-1+w, j0+w)
    __shared__ int sFin[NW][WAVE];       // ... and whether it is final
    const int lane = lane_id();
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int wave = blockIdx.x * NW + wv;
    const int nx = a.nx, ny = a.ny;
    const long np = (long)nx * ny;
    const int n_tiles = a.n_strips_x * a.n_strips_y;
    const StageCoef k = stage_coef(a);
    if (wave == 0 && lane == 0 && a.stage == 0) { a.clock->dt = k.dtd; a.clock->lam[a.parity ^ 1] = 0.0; }
    const float gm = fa.gamma, gm1 = fa.gamma - 1.f, igm1 = 1.f / (fa.gamma - 1.f), gg = gm * igm1;
    const float (&tp)[4][4] = fa.taps;
    const float* U[4] = {(const float*)a.uin[0], (const float*)a.uin[1], (const float*)a.uin[2], (const float*)a.uin[3]};
    const float* U0[4] = {(const float*)a.u0[0], (const float*)a.u0[1], (const float*)a.u0[2], (const float*)a.u0[3]};
    const bool use0 = k.ca != 0.f;   // the combine reads U^n
    float* O[4] = {(float*)a.uout[0], (float*)a.uout[1], (float*)a.uout[2], (float*)a.uout[3]};
    float lam_max = 0.f;
    int bad = 0;
    unsigned long long ntot = 0;

    for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int sxi = t % a.n_strips_x, syi = t / a.n_strips_x;
        const int c = sxi * LIGHT_OUT - 3 + lane;   // lanes 3..60 are the tile's columns
        int cq = c % nx;
        if (cq < 0) cq += nx;
        const bool out_lane = lane >= 3 && lane < WAVE - 3 && c < nx;
        const int cw = c > 0 ? c - 1 : nx - 1;      // west neighbour (output lanes)
        const int j0 = syi * LIGHT_ROWS, nrows = min(LIGHT_ROWS, ny - j0);
        const int r = j0 + wv;                      // the wave's row: face (r-1, r), row r if wv < nrows
        auto row_index = [&](int rr) -> long {
            rr %= ny;
            if (rr < 0) rr += ny;
            return (long)rr * nx + cq;
        };
        float uq[4], u0[4], Fe[4], Fw[4];
        bool finE = false, finW = false;
        // ---- phase A ----
        if (wv <= nrows) {
            // rows r-3..r+2: primitives and sound speed
            float wr[6][4], a0w[6];
            {
                float q[6][4];
#pragma unroll
                for (int m = 0; m < 6; m++) {
                    const long im = row_index(r - 3 + m);
#pragma unroll
                    for (int v = 0; v < 4; v++) q[m][v] = U[v][im];
                }
#pragma unroll
                for (int m = 0; m < 6; m++) {
                    float ir;
                    prim1(q[m], gm1, wr[m], ir);
                    a0w[m] = fast::fsqrt(gm * wr[m][3] * ir);
                }
            }
            if (wv < nrows) {
                const long iu = row_index(r);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    uq[v] = U[v][iu];
                    u0[v] = use0 ? U0[v][iu] : uq[v];
                }
            }
            {
                // face (r-1, r): the north side of row r-1 (rows r-3..r+1 in the frame (rho, v, u, p))
                // and the south side of row r (rows r+2..r-2, oriented -y: (rho, -v, u, p))
                float ws[5][4], fN[4], fS[4], fSp[4], a[4], b[4], ca[4], cb[4], tl[4], tr[4];
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    ws[m][0] = wr[m][0]; ws[m][1] = wr[m][2]; ws[m][2] = wr[m][1]; ws[m][3] = wr[m][3];
                }
                light_value(ws, tp, fN);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    a[v] = ws[2][v];
                    b[v] = ws[3][v];
                    ca[v] = curvature(ws[1][v], ws[2][v], ws[3][v]);
                    cb[v] = curvature(ws[2][v], ws[3][v], ws[4][v]);
                }
                tol_of(wr[2][0], a0w[2], wr[2][3], tl);
                tol_of(wr[3][0], a0w[3], wr[3][3], tr);
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    ws[m][0] = wr[5 - m][0]; ws[m][1] = -wr[5 - m][2]; ws[m][2] = wr[5 - m][1]; ws[m][3] = wr[5 - m][3];
                }
                light_value(ws, tp, fS);
                fSp[0] = fS[0]; fSp[1] = -fS[1]; fSp[2] = fS[2]; fSp[3] = fS[3];
                bool okl, okr;
                face_test(a, b, ca, cb, fN, fSp, tl, tr, okl, okr);
                const float L[4] = {fN[0], fN[2], fN[1], fN[3]};
                const float R[4] = {fS[0], fS[2], -fS[1], fS[3]};
                const fast::C4<float> g = face_flux_t<FLUX, 1>(gm, igm1, gg, L, R);
                sG[wv][lane] = make_float4(g.r, g.ru, g.rv, g.re);
                sFin[wv][lane] = okl && okr ? 1 : 0;
            }
            if (wv < nrows) {
                // x sides of row r from the lane's neighbours -2..+2, and the face east of the lane
                // (its right side and test from lane + 1's values)
                float l1[4], l2[4], r1[4], r2[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    l1[v] = lane_from_left(wr[3][v]);
                    l2[v] = lane_from_left(l1[v]);
                    r1[v] = lane_from_right(wr[3][v]);
                    r2[v] = lane_from_right(r1[v]);
                }
                const float a0r = lane_from_right(a0w[3]);
                float fE[4], fW[4];
                {
                    float ws[5][4];
#pragma unroll
                    for (int v = 0; v < 4; v++) {   // frame (rho, u, v, p), columns -2..+2
                        ws[0][v] = l2[v]; ws[1][v] = l1[v]; ws[2][v] = wr[3][v]; ws[3][v] = r1[v]; ws[4][v] = r2[v];
                    }
                    light_value(ws, tp, fE);
#pragma unroll
                    for (int v = 0; v < 4; v++) {   // oriented -x: (rho, -u, v, p), columns +2..-2
                        const float s0 = v == 1 ? -r2[v] : r2[v], s1 = v == 1 ? -r1[v] : r1[v], s2 = v == 1 ? -wr[3][v] : wr[3][v];
                        const float s3 = v == 1 ? -l1[v] : l1[v], s4 = v == 1 ? -l2[v] : l2[v];
                        ws[0][v] = s0; ws[1][v] = s1; ws[2][v] = s2; ws[3][v] = s3; ws[4][v] = s4;
                    }
                    light_value(ws, tp, fW);
                }
                const float fWr[4] = {lane_from_right(fW[0]), lane_from_right(fW[1]), lane_from_right(fW[2]),
                                      lane_from_right(fW[3])};
                const float R[4] = {fWr[0], -fWr[1], fWr[2], fWr[3]};
                bool okl, okr;
                {
                    float ca[4], cb[4], tl[4], tr[4];
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        ca[v] = curvature(l1[v], wr[3][v], r1[v]);
                        cb[v] = curvature(wr[3][v], r1[v], r2[v]);
                    }
                    tol_of(wr[3][0], a0w[3], wr[3][3], tl);
                    tol_of(r1[0], a0r, r1[3], tr);
                    face_test(wr[3], r1, ca, cb, fE, R, tl, tr, okl, okr);
                }
                const fast::C4<float> f = face_flux_t<FLUX, 0>(gm, igm1, gg, fE, R);
                Fe[0] = f.r; Fe[1] = f.ru; Fe[2] = f.rv; Fe[3] = f.re;
                finE = okl && okr;
#pragma unroll
                for (int v = 0; v < 4; v++) Fw[v] = lane_from_left(Fe[v]);
                finW = lane_from_left_i(finE ? 1 : 0) != 0;
            }
        }
        __syncthreads();
        // ---- phase B: the update of row r from the faces (r-1, r) and (r, r+1) ----
        if (wv < nrows) {
            float Gs[4], Gn[4];
            const float4 gs = sG[wv][lane], gn = sG[wv + 1][lane];
            Gs[0] = gs.x; Gs[1] = gs.y; Gs[2] = gs.z; Gs[3] = gs.w;
            Gn[0] = gn.x; Gn[1] = gn.y; Gn[2] = gn.z; Gn[3] = gn.w;
            const bool finS = sFin[wv][lane] != 0, finN = sFin[wv + 1][lane] != 0;
            const long qc = (long)r * nx + c;
            const unsigned mask = (finW ? 0u : 1u) | (finE ? 0u : 2u) | (finS ? 0u : 4u) | (finN ? 0u : 8u);
            const bool pend = out_lane && mask != 0u;
            ntot += __popcll(__ballot(out_lane && finE)) + __popcll(__ballot(out_lane && finN));
            if (out_lane && !pend) {
                float o[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    o[v] = cell_update(k, Fw[v], Fe[v], Gs[v], Gn[v], uq[v], u0[v]);
                    O[v][qc] = o[v];
                }
                if (k.fuse) {
                    float l;
                    if (lam_cell_f32<false>(a, o[0], o[1], o[2], o[3], l)) lam_max = fmaxf(lam_max, l);
                    else bad++;
                }
            }
            const unsigned long long pm = __ballot(pend);
            if (pm) {
                // the cell waits for its deferred faces: list it with their mask and leave the
                // fluxes of its final faces in the planes
                unsigned base = 0u;
                if (lane == 0) base = atomicAdd(fa.n_list, (unsigned)__popcll(pm));
                base = __shfl(base, 0, WAVE);
                if (pend) {
                    fa.list[base + (unsigned)__popcll(pm & ((1ull << lane) - 1ull))] = ((unsigned)qc << 4) | mask;
                    const long qw = (long)r * nx + cw, qs = (long)(r > 0 ? r - 1 : ny - 1) * nx + c;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        if (finW) fa.F[v * np + qw] = Fw[v];
                        if (finE) fa.F[v * np + qc] = Fe[v];
                        if (finS) fa.F[(4 + v) * np + qs] = Gs[v];
                        if (finN) fa.F[(4 + v) * np + qc] = Gn[v];
                    }
                }
            }
        }
        __syncthreads();   // the next tile rewrites the flux slots
    }
    if (fa.stats && lane == 0 && ntot) fa.stats[3 * wave + 1] += ntot;   // the wave's own slot
    if (wave == 0 && lane == 0 && (a.euler || a.stage == 2)) {
        a.clock->time += k.dtd;
        a.clock->step += 1;
    }
    if (k.fuse) {
        const double wm = wave_max((double)lam_max);
        const long long b = wave_sum_ll((long long)bad);
        if (lane == 0) {
            a.lam_part[wave] = wm;   // one slot per wave, reduced by lam_reduce_kernel
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
}

// The cells stage_light_kernel left (and, on the last stage, the fold of its CFL slots): per
// list entry (cell << 4 | mask of deferred faces,
// bits west, east, south, north) the deferred faces through the network as the deferred-face
// pass evaluates them (net_face, face_flux), the others from the flux planes, then the update
// and the CFL.  A face deferred between two listed cells is evaluated for each, with the same
// operands and instructions, so both see the same flux.  Persistent grid; the last workgroup
// to finish clears the list count (n_list[1] counts finished workgroups), so the next stage
// starts from an empty list and the sequence can be captured in a graph.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void pending_kernel(FaceArgs fa, StageArgs a) {
    const unsigned n = *fa.n_list;
    const bool fuse = a.fuse_cfl && (a.euler || a.stage == 2);
    if (fuse && fa.n_lam_slots > 0) {
        // the stage kernel's per-wave CFL maxima -> lam[parity ^ 1], in place of a reduction
        // launch: each workgroup folds its share of the slots (a few per thread) and takes an
        // atomic maximum, as this kernel's own cells' maxima do below
        __shared__ double red[4];
        const int per = (fa.n_lam_slots + (int)gridDim.x - 1) / (int)gridDim.x;
        const int i0 = (int)blockIdx.x * per, i1 = min(i0 + per, fa.n_lam_slots);
        double m = 0.0;
        for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) m = fmax(m, a.lam_part[i]);
        m = wave_max(m);
        if (lane_id() == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
            if (m > 0.0) atomic_max_pos_double(&a.clock->lam[a.parity ^ 1], m);
        }
    }
    if (n == 0u) return;    // no face deferred (the usual case on smooth flow)
    __shared__ __attribute__((aligned(16))) uint32_t lds[4][128][SLOT];
    __shared__ float skip[SK_N];
    stage_skip(fa.w, skip);
    init_feat_slots(lds[threadIdx.x >> 6]);
    Regs R;
    load_regs(fa.w, lane_id(), R);
    uint32_t (*feat)[SLOT] = lds[threadIdx.x >> 6];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int nx = fa.nx, ny = fa.ny;
    const long np = (long)nx * ny;
    const bool own4 = fa.w->own4 != 0;
    const StageCoef k = stage_coef(a);
    const bool use0 = k.ca != 0.f;
    unsigned long long nfb = 0, ntot = 0, nnet = 0;
    float lam_max = 0.f;
    int bad = 0;
    const int n_tasks = (int)((n + 63u) / 64u);
    for (int task = blockIdx.x * 4 + wv; task < n_tasks; task += gridDim.x * 4) {
        const unsigned e = (unsigned)task * 64u + (unsigned)lane;
        const bool valid = e < n;
        const unsigned ent = valid ? fa.list[e] : 0u;
        const long qc = (long)(ent >> 4);
        const unsigned mask = ent & 15u;
        const int j = (int)(qc / nx), i = (int)(qc - (long)j * nx);
        const int iw = i > 0 ? i - 1 : nx - 1, js = j > 0 ? j - 1 : ny - 1;
        float Fw[4] = {0.f, 0.f, 0.f, 0.f}, Fe[4] = {0.f, 0.f, 0.f, 0.f}, Gs[4] = {0.f, 0.f, 0.f, 0.f},
              Gn[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int d = 0; d < 4; d++) {   // west, east, south, north
            const int axis = d >> 1;
            const int row = d == 2 ? js : j, cc = d == 0 ? iw : i;
            const bool need = valid && ((mask >> d) & 1u) != 0u;
            float f[4] = {0.f, 0.f, 0.f, 0.f};
            if (__ballot(need)) {
                float L[4], Rr[4];
                bool fell, need0, need1;
                net_face(fa, skip, feat, R, true, own4, axis, row, cc, need, L, Rr, fell, need0, need1);
                const bool own = need && (d & 1) != 0;   // the east and north faces are the cell's own
                nnet += __popcll(__ballot(own && need0)) + __popcll(__ballot(own && need1));
                nfb += __popcll(__ballot(own && fell));
                ntot += __popcll(__ballot(own));
                const fast::C4<float> ff = face_flux(fa, axis, L, Rr);
                f[0] = ff.r; f[1] = ff.ru; f[2] = ff.rv; f[3] = ff.re;
            }
            if (valid && !need) {
                const long fi = (long)row * nx + cc;
#pragma unroll
                for (int v = 0; v < 4; v++) f[v] = fa.F[(axis * 4 + v) * np + fi];
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (d == 0) Fw[v] = f[v];
                else if (d == 1) Fe[v] = f[v];
                else if (d == 2) Gs[v] = f[v];
                else Gn[v] = f[v];
            }
        }
        if (valid) {
            float o[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float u = ((const float*)a.uin[v])[qc];
                const float u0 = use0 ? ((const float*)a.u0[v])[qc] : u;
                o[v] = cell_update(k, Fw[v], Fe[v], Gs[v], Gn[v], u, u0);
                ((float*)a.uout[v])[qc] = o[v];
            }
            if (k.fuse) {
                float l;
                if (lam_cell_f32<false>(a, o[0], o[1], o[2], o[3], l)) lam_max = fmaxf(lam_max, l);
                else bad++;
            }
        }
    }
    if (fa.stats && lane == 0 && (ntot || nnet)) {
        unsigned long long* st = fa.stats + 3 * (blockIdx.x * 4 + wv);
        st[0] += nfb;
        st[1] += ntot;
        st[2] += nnet;
    }
    if (k.fuse) {
        const double w = wave_max((double)lam_max);
        const long long b = wave_sum_ll((long long)bad);
        if (lane == 0) {
            if (w > 0.0) atomic_max_pos_double(&a.clock->lam[a.parity ^ 1], w);
            if (b) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)b);
        }
    }
    // every thread of the block has read the count; the last block clears it
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned t = atomicAdd(fa.n_list + 1, 1u);
        if (t == gridDim.x - 1u) {
            fa.n_list[0] = 0u;
            fa.n_list[1] = 0u;
        }
    }
}

// The matrix-core MLP alone on n given feature rows X (n x 20): Y = base + skip + MLP.
__global__ __launch_bounds__(256) void infer_kernel(const Packed* __restrict__ P, const float* __restrict__ X, long n,
                                                    float* __restrict__ Y) {
    __shared__ __attribute__((aligned(16))) uint32_t lds[4][64][SLOT];
    __shared__ float skip[SK_N];
    stage_skip(P, skip);
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int g = lane >> 4, jl = lane & 15;
    uint32_t(*feat)[SLOT] = lds[wv];
    for (int d = 10; d < SLOT; d++) feat[lane][d] = 0u;
    Regs R;
    load_regs(P, lane, R);
    for (long base = (long)blockIdx.x * 256; base < n; base += (long)gridDim.x * 256) {
        const long i = base + wv * 64 + lane;
        const long ii = i < n ? i : n - 1;
        float x[20];
#pragma unroll
        for (int f = 0; f < 20; f++) x[f] = X[ii * 20 + f];
        float yb[4], sk[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float s = skip[SK_BL + v];
#pragma unroll
            for (int f = 0; f < 20; f++) s = fmaf(skip[SK_L + v * NIN + f], x[f], s);
            yb[v] = 0.25f * (x[12 + v] - x[4 + v]);
            sk[v] = s;
        }
        uint32_t* slot = feat[lane];
        *(u32x4*)(slot) = u32x4{pack2(x[0], x[1]), pack2(x[2], x[3]), pack2(x[4], x[5]), pack2(x[6], x[7])};
        *(u32x4*)(slot + 4) = u32x4{pack2(x[8], x[9]), pack2(x[10], x[11]), pack2(x[12], x[13]), pack2(x[14], x[15])};
        slot[8] = pack2(x[16], x[17]);
        slot[9] = pack2(x[18], x[19]);
        slot[10] = pack2(1.f, 0.f);
        __syncthreads();
        f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int qd = 0; qd < 4; qd++) {
            const bf16x8 bx = __builtin_bit_cast(bf16x8, *(const u32x4*)&feat[qd * 16 + jl][4 * g]);
            const f32x4 d3 = mlp16(R, bx);
            if (g == qd) o = d3;
        }
        __syncthreads();
        if (i < n) {
#pragma unroll
            for (int v = 0; v < 4; v++) Y[i * 4 + v] = fmaf(gate(o[v] + skip[SK_B3 + v]), sk[v], yb[v]);
        }
    }
}

// ---------------------------------------------------------------------------
// training (FP32).  One block = one wave = 64 samples, one thread per sample.
// The weights and every per-sample vector live in LDS (rows of stride 65, so a
// thread's own row hits a different bank from its neighbours'), the loops stay
// rolled and the registers few.  Weight gradients are sums over the block's
// samples, written as per-block partials and reduced in a fixed order.
// ---------------------------------------------------------------------------
constexpr int RS = NH + 1;

__device__ __forceinline__ float gelu_d(float x) {
    const float t = tanhf(GELU_C * (x + GELU_A * x * x * x));
    return 0.5f * x * (1.f + t);
}
__device__ __forceinline__ float gelu_grad_d(float x) {
    const float t = tanhf(GELU_C * (x + GELU_A * x * x * x));
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * GELU_C * (1.f + 3.f * GELU_A * x * x);
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(TB) void train_kernel(TrainArgs a) {
    __shared__ float sW1[NH][NIN];
    __shared__ float sW2[NH][NH];
    __shared__ float sW3[NOUT][NH];
    __shared__ float sL[NOUT][NIN];
    __shared__ float sB1[NH], sB2[NH], sB3[NOUT], sBL[NOUT];
    __shared__ float sX[TB][NIN + 1];
    __shared__ float sZ1[TB][RS];   // z1, then dz1
    __shared__ float sH1[TB][RS];
    __shared__ float sZ2[TB][RS];   // z2, then dz2
    __shared__ float sH2[TB][RS];
    __shared__ float sDz3[TB][NOUT];     // d loss / d (gate pre-activation)
    __shared__ float sDs[TB][NOUT];      // d loss / d (learned correction) = dy * phi
    __shared__ double sLoss[TB];
    const int t = threadIdx.x;
    const float* P = a.params;
    for (int i = t; i < NH * NIN; i += TB) sW1[i / NIN][i % NIN] = P[OFF_W1 + i];
    for (int i = t; i < NH * NH; i += TB) sW2[i / NH][i % NH] = P[OFF_W2 + i];
    for (int i = t; i < NOUT * NH; i += TB) sW3[i / NH][i % NH] = P[OFF_W3 + i];
    for (int i = t; i < NOUT * NIN; i += TB) sL[i / NIN][i % NIN] = P[OFF_L + i];
    for (int i = t; i < NH; i += TB) { sB1[i] = P[OFF_B1 + i]; sB2[i] = P[OFF_B2 + i]; }
    if (t < NOUT) { sB3[t] = P[OFF_B3 + t]; sBL[t] = P[OFF_BL + t]; }
    const long bi = (long)blockIdx.x * TB + t;
    const bool live = bi < a.n;
    long si = 0;
    if (live)
        si = a.random ? (long)(mix64(a.seed ^ mix64(a.step * 0x100000001B3ull + (unsigned long long)bi)) % (unsigned long long)a.n_data)
                      : bi;
    float x[NIN];
#pragma unroll
    for (int f = 0; f < NIN; f++) { x[f] = live ? a.X[si * NIN + f] : 0.f; sX[t][f] = x[f]; }
    __syncthreads();
    // forward
    for (int o = 0; o < NH; o++) {
        float z = sB1[o];
#pragma unroll
        for (int f = 0; f < NIN; f++) z = fmaf(sW1[o][f], x[f], z);
        sZ1[t][o] = z;
        sH1[t][o] = gelu_d(z);
    }
    float z3[NOUT], sk[NOUT];        // gate pre-activations, learned correction L x + bL
#pragma unroll
    for (int v = 0; v < NOUT; v++) {
        float s = sBL[v];
#pragma unroll
        for (int f = 0; f < NIN; f++) s = fmaf(sL[v][f], x[f], s);
        sk[v] = s;
        z3[v] = sB3[v];
    }
    for (int o = 0; o < NH; o++) {
        float z = sB2[o];
#pragma unroll 8
        for (int j = 0; j < NH; j++) z = fmaf(sW2[o][j], sH1[t][j], z);
        sZ2[t][o] = z;
        const float h = gelu_d(z);
        sH2[t][o] = h;
#pragma unroll
        for (int v = 0; v < NOUT; v++) z3[v] = fmaf(sW3[v][o], h, z3[v]);
    }
    float dz3[NOUT];
    double l2 = 0.0;
#pragma unroll
    for (int v = 0; v < NOUT; v++) {
        const float phi = 1.f / (1.f + expf(-z3[v]));
        const float y = 0.25f * (x[12 + v] - x[4 + v]) + phi * sk[v];
        const float r = live ? y - a.Y[si * NOUT + v] : 0.f;
        const float wv = sample_weight(x[8 + v]);
        l2 += (double)wv * (double)r * (double)r;
        const float dy = a.scale * wv * r;
        dz3[v] = dy * sk[v] * phi * (1.f - phi);
        sDz3[t][v] = dz3[v];
        sDs[t][v] = dy * phi;
    }
    sLoss[t] = l2;
    // backward through layer 3 and 2 (per sample)
    for (int j = 0; j < NH; j++) {
        float d = 0.f;
#pragma unroll
        for (int v = 0; v < NOUT; v++) d = fmaf(sW3[v][j], dz3[v], d);
        sZ2[t][j] = d * gelu_grad_d(sZ2[t][j]);
    }
    for (int j = 0; j < NH; j++) {
        float d = 0.f;
#pragma unroll 8
        for (int o = 0; o < NH; o++) d = fmaf(sW2[o][j], sZ2[t][o], d);
        sZ1[t][j] = d * gelu_grad_d(sZ1[t][j]);
    }
    __syncthreads();
    float* part = a.partial + (long)blockIdx.x * NPARAM;
    // dW2 row t and db2[t]: sum over samples of dz2[s][t] * h1[s][:]
    {
        float acc[NH];
#pragma unroll
        for (int j = 0; j < NH; j++) acc[j] = 0.f;
        float bsum = 0.f;
        for (int s = 0; s < TB; s++) {
            const float d = sZ2[s][t];
            bsum += d;
#pragma unroll
            for (int j = 0; j < NH; j++) acc[j] = fmaf(d, sH1[s][j], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < NH; j++) part[OFF_W2 + t * NH + j] = acc[j];
        part[OFF_B2 + t] = bsum;
    }
    // dW1 row t and db1[t]
    {
        float acc[NIN];
#pragma unroll
        for (int f = 0; f < NIN; f++) acc[f] = 0.f;
        float bsum = 0.f;
        for (int s = 0; s < TB; s++) {
            const float d = sZ1[s][t];
            bsum += d;
#pragma unroll
            for (int f = 0; f < NIN; f++) acc[f] = fmaf(d, sX[s][f], acc[f]);
        }
#pragma unroll
        for (int f = 0; f < NIN; f++) part[OFF_W1 + t * NIN + f] = acc[f];
        part[OFF_B1 + t] = bsum;
    }
    // dW3[v][t], db3 (through the gate); dL, dbL (through the gated correction)
#pragma unroll
    for (int v = 0; v < NOUT; v++) {
        float acc = 0.f;
        for (int s = 0; s < TB; s++) acc = fmaf(sDz3[s][v], sH2[s][t], acc);
        part[OFF_W3 + v * NH + t] = acc;
    }
    for (int i = t; i < NOUT * NIN + 2 * NOUT; i += TB) {
        float acc = 0.f;
        if (i < NOUT * NIN) {
            const int v = i / NIN, f = i % NIN;
            for (int s = 0; s < TB; s++) acc = fmaf(sDs[s][v], sX[s][f], acc);
            part[OFF_L + i] = acc;
        } else if (i < NOUT * NIN + NOUT) {
            const int v = i - NOUT * NIN;
            for (int s = 0; s < TB; s++) acc += sDs[s][v];
            part[OFF_BL + v] = acc;
        } else {
            const int v = i - NOUT * NIN - NOUT;
            for (int s = 0; s < TB; s++) acc += sDz3[s][v];
            part[OFF_B3 + v] = acc;
        }
    }
    if (t == 0) {
        double l = 0.0;
        for (int s = 0; s < TB; s++) l += sLoss[s];
        a.loss_partial[blockIdx.x] = l;
    }
}

// Weighted normal equations of the linear skip, per variable v:
//     G_v = sum_s w_sv a_s a_s^T,   b_v = sum_s w_sv a_s (y*_sv - base_sv),   a_s = (x_s, 1),
// in FP64 over every `stride`-th sample.  Entry e of the packed [v][NA][NA + 1] array
// (Gram columns, then the right-hand side) is owned by thread e % 256 of each block; a
// block sums its tiles of 64 samples and writes one partial row.
constexpr int NA = NIN + 1;
constexpr int NGRAM = NOUT * NA * (NA + 1);
constexpr int GRAM_PER_THREAD = (NGRAM + 255) / 256;

__global__ __launch_bounds__(256) void gram_kernel(const float* __restrict__ X, const float* __restrict__ Y, long n,
                                                   long stride, double* __restrict__ partial) {
    __shared__ float sa[64][NA + 1];
    __shared__ float sr[64][NOUT];
    __shared__ float sw[64][NOUT];
    double acc[GRAM_PER_THREAD];
#pragma unroll
    for (int k = 0; k < GRAM_PER_THREAD; k++) acc[k] = 0.0;
    const long n_tiles = (n + 64 * stride - 1) / (64 * stride);
    for (long tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        if (threadIdx.x < 64) {
            const int i = threadIdx.x;
            const long s = (tile * 64 + i) * stride;
            const bool ok = s < n;
            float x[NIN];
#pragma unroll
            for (int f = 0; f < NIN; f++) { x[f] = ok ? X[s * NIN + f] : 0.f; sa[i][f] = x[f]; }
            sa[i][NIN] = 1.f;
#pragma unroll
            for (int v = 0; v < NOUT; v++) {
                sr[i][v] = ok ? Y[s * NOUT + v] - 0.25f * (x[12 + v] - x[4 + v]) : 0.f;
                sw[i][v] = ok ? sample_weight(x[8 + v]) : 0.f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < GRAM_PER_THREAD; k++) {
            const int e = threadIdx.x + 256 * k;
            if (e < NGRAM) {
                const int v = e / (NA * (NA + 1)), rem = e % (NA * (NA + 1));
                const int i = rem / (NA + 1), j = rem % (NA + 1);
                double s = 0.0;
                for (int q = 0; q < 64; q++) {
                    const float other = j < NA ? sa[q][j] : sr[q][v];
                    s += (double)(sw[q][v] * sa[q][i]) * (double)other;
                }
                acc[k] += s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < GRAM_PER_THREAD; k++) {
        const int e = threadIdx.x + 256 * k;
        if (e < NGRAM) partial[(long)blockIdx.x * NGRAM + e] = acc[k];
    }
}

// grad[p] = sum over blocks of partial[b][p] (fixed order); loss = sum of block losses / (n NOUT)
__global__ void reduce_kernel(const float* __restrict__ partial, int nblk, float* __restrict__ grad,
                              const double* __restrict__ loss_partial, double inv, double* loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPARAM) {
        float acc = 0.f;
        for (int b = 0; b < nblk; b++) acc += partial[(long)b * NPARAM + p];
        grad[p] = acc;
    }
    if (p == 0 && loss_out) {
        double l = 0.0;
        for (int b = 0; b < nblk; b++) l += loss_partial[b];
        *loss_out = l * inv;
    }
}

// Adam (Kingma & Ba 2015, Algorithm 1), step t >= 1
// over the first n_update parameters (NPARAM, or OFF_L to keep the linear skip fixed)
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                            float lr, float b1, float b2, float eps, float c1, float c2, int n_update) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_update) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ---------------------------------------------------------------------------
// training samples from a fine state (FP64 arithmetic; fine planes FP32 or FP64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ldf(const SampleArgs& a, int v, long i) {
    return a.esz == 8 ? ((const double*)a.U[v])[i] : (double)((const float*)a.U[v])[i];
}

// box filter f x f -> coarse cell averages, or the coarse solver's own cells when given
// (features of the states the network meets inside the solver, targets from the fine run)
__global__ void coarse_kernel(SampleArgs a) {
    const long nc = (long)a.nc_x * a.nc_y;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int I = (int)(c % a.nc_x), J = (int)(c / a.nc_x);
    if (a.Uc[0]) {
        for (int v = 0; v < 4; v++)
            a.cells[v * nc + c] = a.esz_c == 8 ? ((const double*)a.Uc[v])[c] : (double)((const float*)a.Uc[v])[c];
        return;
    }
    for (int v = 0; v < 4; v++) {
        double s = 0.0;
        for (int b = 0; b < a.f; b++)
            for (int e = 0; e < a.f; e++) s += ldf(a, v, (long)(J * a.f + b) * a.nxf + (I * a.f + e));
        a.cells[v * nc + c] = s / (double)(a.f * a.f);
    }
}

// Line averages of the fine solution over the coarse east / north faces: per fine row
// (column) the fourth-order face value (-u[m-1] + 7u[m] + 7u[m+1] - u[m+2]) / 12.
__global__ void face_target_kernel(SampleArgs a) {
    const long nc = (long)a.nc_x * a.nc_y;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= 2 * nc) return;
    const int axis = c >= nc ? 1 : 0;
    const long cc = c - axis * nc;
    const int I = (int)(cc % a.nc_x), J = (int)(cc / a.nc_x);
    for (int v = 0; v < 4; v++) {
        double s = 0.0;
        for (int b = 0; b < a.f; b++) {
            long i[4];
            for (int k = 0; k < 4; k++) {
                if (axis == 0) {
                    const int m = wrap(I * a.f + a.f - 2 + k, a.nxf);
                    i[k] = (long)(J * a.f + b) * a.nxf + m;
                } else {
                    const int m = wrap(J * a.f + a.f - 2 + k, a.nyf);
                    i[k] = (long)m * a.nxf + (I * a.f + b);
                }
            }
            s += (-ldf(a, v, i[0]) + 7.0 * ldf(a, v, i[1]) + 7.0 * ldf(a, v, i[2]) - ldf(a, v, i[3])) / 12.0;
        }
        a.faces[(axis * 4 + v) * nc + cc] = s / (double)a.f;
    }
}

__device__ __forceinline__ void prim_d(const double q[4], double gm1, double w[4]) {
    w[0] = q[0];
    w[1] = q[1] / q[0];
    w[2] = q[2] / q[0];
    w[3] = gm1 * (q[3] - 0.5 * (q[1] * w[1] + q[2] * w[2]));
}

// one sample per (axis, side, coarse cell), in that order
__global__ void sample_kernel(SampleArgs a, long n_out) {
    const long nc = (long)a.nc_x * a.nc_y;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_out) return;
    const int axis = (int)(id / (2 * nc));
    const int sd = (int)((id / nc) & 1);
    const long cell = id % nc;
    const int I = (int)(cell % a.nc_x), J = (int)(cell / a.nc_x);
    const int sgn = sd == 0 ? 1 : -1;
    const double gm1 = a.gamma - 1.0;
    double ws[5][4];
    for (int k = 0; k < 5; k++) {
        const int off = sgn * (k - 2);
        const long ci = axis == 0 ? (long)J * a.nc_x + wrap(I + off, a.nc_x) : (long)wrap(J + off, a.nc_y) * a.nc_x + I;
        double q[4], w[4];
        for (int v = 0; v < 4; v++) q[v] = a.cells[v * nc + ci];
        prim_d(q, gm1, w);
        const double un = axis == 0 ? w[1] : w[2], ut = axis == 0 ? w[2] : w[1];
        ws[k][0] = w[0]; ws[k][1] = sgn * un; ws[k][2] = ut; ws[k][3] = w[3];
    }
    // target: the face between the cell and its +sgn neighbour
    long fi = cell;
    if (sgn < 0) fi = axis == 0 ? (long)J * a.nc_x + wrap(I - 1, a.nc_x) : (long)wrap(J - 1, a.nc_y) * a.nc_x + I;
    double qf[4], wf[4];
    for (int v = 0; v < 4; v++) qf[v] = a.faces[(axis * 4 + v) * nc + fi];
    prim_d(qf, gm1, wf);
    const double tun = axis == 0 ? wf[1] : wf[2], tut = axis == 0 ? wf[2] : wf[1];
    const double to[4] = {wf[0], sgn * tun, tut, wf[3]};
    const double a0 = sqrt(a.gamma * ws[2][3] / ws[2][0]);
    const double scale[4] = {ws[2][0], a0, a0, ws[2][3]};
    float* X = a.X + id * NIN;
    float* Y = a.Y + id * NOUT;
    for (int v = 0; v < 4; v++) {
        const double c = ws[2][v];
        double m = (double)FLOOR * scale[v];
        for (int k = 0; k < 5; k++)
            if (k != 2) m = fmax(m, fabs(ws[k][v] - c));
        for (int k = 0; k < 5; k++)
            if (k != 2) X[4 * k + v] = (float)((ws[k][v] - c) / m);
        X[8 + v] = (float)(m / (m + (double)X0C * scale[v]));
        Y[v] = (float)fmin(4.0, fmax(-4.0, (to[v] - c) / m));
    }
}

}  // namespace mlp

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t launch_mlp_faces(const mlp::FaceArgs& a, hipStream_t st) {
    // persistent: two 4-wave blocks per CU (the kernel's register budget allows 2 waves per
    // SIMD); each wave loads the network into registers once and loops over tasks
    const int blocks = std::min((a.n_tasks + 3) / 4, 512);
    mlp::faces_kernel<<<blocks, 256, 0, st>>>(a);
    return hipGetLastError();
}

// Bypass stage: light pass over every face, then the network over the deferred ones.
// The list length stays on the device (the list pass's persistent grid reads it), so
// the sequence can be captured in a graph; the stage kernel that consumes the face
// states clears the counter for the next stage (it starts at zero when allocated).
hipError_t launch_mlp_faces_bypass(const mlp::FaceArgs& a, hipStream_t st) {
    // one task (64 faces) per wave, at most STAT_SLOTS waves: the 4-tap light pass holds 67
    // VGPRs, so 7 of its waves fit per SIMD and a 512^2 stage's 8192 waves are nearly all in
    // flight at once (two tasks per wave measured 2% slower).  The list pass is persistent and
    // small: deferred faces are rare.
    const int lb = std::min((a.n_tasks + 3) / 4, mlp::STAT_SLOTS / 4);
    if (a.own4) mlp::faces_light_kernel<true><<<lb, 256, 0, st>>>(a);
    else mlp::faces_light_kernel<false><<<lb, 256, 0, st>>>(a);
    mlp::faces_list_kernel<<<128, 256, 0, st>>>(a);
    return hipGetLastError();
}

// Fused bypass stage: stage_light_kernel over a.n_strips_x x a.n_strips_y tiles of LIGHT_OUT
// columns by LIGHT_ROWS rows on `blocks` workgroups of LIGHT_ROWS + 1 waves (at most STAT_SLOTS
// waves: one statistics slot and one CFL slot each), then pending_kernel for the cells it
// listed.  fa.list / fa.n_list: the pending cells and their count, then pending_kernel's
// finished-workgroup counter (both zero between stages).
hipError_t launch_mlp_light_stage(const StageArgs& a, const mlp::FaceArgs& fa, int blocks, hipStream_t st) {
    switch (fa.flux) {
    case 0: mlp::stage_light_kernel<0><<<blocks, 64 * (mlp::LIGHT_ROWS + 1), 0, st>>>(a, fa); break;
    case 1: mlp::stage_light_kernel<1><<<blocks, 64 * (mlp::LIGHT_ROWS + 1), 0, st>>>(a, fa); break;
    default: mlp::stage_light_kernel<2><<<blocks, 64 * (mlp::LIGHT_ROWS + 1), 0, st>>>(a, fa); break;
    }
    mlp::pending_kernel<<<128, 256, 0, st>>>(fa, a);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const mlp::Packed* P, const float* X, long n, float* Y, hipStream_t st) {
    const long blocks = std::min<long>((n + 255) / 256, 4096);
    mlp::infer_kernel<<<(int)blocks, 256, 0, st>>>(P, X, n, Y);
    return hipGetLastError();
}

hipError_t launch_mlp_train(const mlp::TrainArgs& a, int nblk, float* grad, double* loss_out, hipStream_t st) {
    mlp::train_kernel<<<nblk, mlp::TB, 0, st>>>(a);
    mlp::reduce_kernel<<<(mlp::NPARAM + 255) / 256, 256, 0, st>>>(a.partial, nblk, grad, a.loss_partial,
                                                                  1.0 / ((double)a.n * mlp::NOUT), loss_out);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* p, const float* g, float* m, float* v, float lr, int t, int n_update, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, t))), c2 = (float)(1.0 / (1.0 - std::pow(0.999, t)));
    mlp::adam_kernel<<<(mlp::NPARAM + 255) / 256, 256, 0, st>>>(p, g, m, v, lr, b1, b2, 1e-8f, c1, c2, n_update);
    return hipGetLastError();
}

hipError_t launch_mlp_gram(const float* X, const float* Y, long n, long stride, double* partial, int blocks,
                           hipStream_t st) {
    mlp::gram_kernel<<<blocks, 256, 0, st>>>(X, Y, n, stride, partial);
    return hipGetLastError();
}

hipError_t launch_mlp_samples(const mlp::SampleArgs& a, long n_out, hipStream_t st) {
    const long nc = (long)a.nc_x * a.nc_y;
    mlp::coarse_kernel<<<(int)((nc + 255) / 256), 256, 0, st>>>(a);
    mlp::face_target_kernel<<<(int)((2 * nc + 255) / 256), 256, 0, st>>>(a);
    mlp::sample_kernel<<<(int)((n_out + 255) / 256), 256, 0, st>>>(a, n_out);
    return hipGetLastError();
}

}  // namespace efvm
