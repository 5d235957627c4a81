// efvm_mlp.hip -- learned face reconstruction on gfx950.
//
//  * faces_kernel: for every face of a periodic grid, the two face sides' stencil
//    features, the MLP on bf16 16x16x32 matrix cores (FP32 accumulation), the
//    FP32 skip and offsets, the per-face positivity / TVD check with fallback to
//    the solver's MUSCL (conserved variables, MC limiter), and the final face
//    primitives for the learned stage kernel.
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

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Write one face: its primitives (out_mode 0) or its Riemann flux (out_mode 1).  L, R are
// the lab-frame primitives (rho, u, v, p) of the left / lower and right / upper sides.
__device__ __forceinline__ void put_face(const FaceArgs& a, int axis, long cell, long np, const float L[4], const float R[4]) {
    if (a.out_mode == 1) {
        const float gm = a.gamma, igm1 = 1.f / (a.gamma - 1.f), gg = gm * igm1;
        const fast::Prep pl = fast::prep_prim(L[0], L[1], L[2], L[3], gm, igm1);
        const fast::Prep pr = fast::prep_prim(R[0], R[1], R[2], R[3], gm, igm1);
        fast::C4<float> f;
        if (axis == 0) f = a.flux == 0 ? fast::flux<0, 0>(pl, pr, gg) : (a.flux == 1 ? fast::flux<1, 0>(pl, pr, gg) : fast::flux<2, 0>(pl, pr, gg));
        else f = a.flux == 0 ? fast::flux<0, 1>(pl, pr, gg) : (a.flux == 1 ? fast::flux<1, 1>(pl, pr, gg) : fast::flux<2, 1>(pl, pr, gg));
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
__device__ __forceinline__ void to_prim(const float q[6][4], float gm1, float w[6][4]) {
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const float ir = fast::rcp(q[m][0]);
        w[m][0] = q[m][0];
        w[m][1] = q[m][1] * ir;
        w[m][2] = q[m][2] * ir;
        w[m][3] = gm1 * (q[m][3] - 0.5f * (q[m][1] * w[m][1] + q[m][2] * w[m][2]));
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

template <int PASS>   // PASS_ALL or PASS_LIST (PASS_LIGHT is faces_light_kernel)
__device__ __forceinline__ void faces_body(const FaceArgs& a, const float* skip, uint32_t (*feat)[SLOT], const Regs& R) {
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int g = lane >> 4, jl = lane & 15;
    const float gm = a.gamma, gm1 = a.gamma - 1.f;
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
        float sig[2][4], cen[2][4], y[2][4], sk[2][4], tol[2][4];
        bool need0 = true, need1 = true;   // the side's gate is evaluated by the network
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
        const long cell = (long)row * a.nx + cc;
        float fo[2][4];
        bool ok = true;
        float q[6][4], w[6][4];
        {
            wave_sync();
            const unsigned long long nm0 = __ballot(valid && need0), nm1 = __ballot(valid && need1);
            nnet += __popcll(nm0) + __popcll(nm1);
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
        float L[4], Rr[4];   // lab-frame primitives of the left / right face states
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
        nfb += __popcll(__ballot(valid && !ok));
        ntot += __popcll(__ballot(valid));
        if (valid) put_face(a, axis, cell, np, L, Rr);
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
    unsigned long long ntot = 0;
    for (int task = blockIdx.x * 4 + wv; task < a.n_tasks; task += gridDim.x * 4) {
        int axis, row, cc;
        bool valid;
        task_face(a, task, lane, axis, row, cc, valid);
        float q[6][4], w[6][4], fo[2][4];
        load_cells(a, axis, row, cc, q, w);
        bool defer = false;
#pragma unroll
        for (int sd = 0; sd < 2; sd++) {
            float ws[5][4], tol[4];
            orient(w, axis, sd, ws);
            if (own4) {
                // gate open, 4-tap correction, no bias: the normalisation cancels and the face
                // value is the linear stencil w0 + sum_k t_k (w_k - w0) with t = L + central slope
                const float a0 = fast::fsqrt(gm * ws[2][3] * fast::rcp(ws[2][0]));
                const float scale[4] = {ws[2][0], a0, a0, ws[2][3]};
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float c = ws[2][v];
                    const float* Lv = skip + SK_L + v * NIN;
                    float d = Lv[v] * (ws[0][v] - c);
                    d = fmaf(Lv[4 + v] - 0.25f, ws[1][v] - c, d);
                    d = fmaf(Lv[12 + v] + 0.25f, ws[3][v] - c, d);
                    d = fmaf(Lv[16 + v], ws[4][v] - c, d);
                    fo[sd][v] = c + d;
                    tol[v] = FLOOR * scale[v];
                }
            } else {
                float sig[4], cen[4], y[4], sk[4];
                featurize<false>(ws, gm, skip, own4, nullptr, sig, cen, y, sk, tol);
#pragma unroll
                for (int v = 0; v < 4; v++) fo[sd][v] = fmaf(sig[v], y[v] + sk[v], cen[v]);   // gate fully open
            }
            defer = defer || !side_ok(ws, fo[sd], tol, 0.25f);
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
