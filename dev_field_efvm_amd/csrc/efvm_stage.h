// efvm_stage.h -- the fused RK stage kernels (HIP, gfx950).
//
// One launch = one RK substage of the reference's phase sequence
// (crates/cfd_solver/src/lib.rs:919-931; examples: one app.run()):
//   periodic/ghost/boundary fill  ->  limited slopes  ->  face states  ->
//   Riemann flux  ->  divergence (+ viscous, body force, solid-rhs zero)  ->
//   SSP-RK3 combine  [-> fused CFL lambda-max of the new state, stage 3].
//
// Layout: conserved variables in four SoA planes (rho, rho_u, rho_v, rho_E) of
// nx*ny interior cells, x fastest; rho_w == 0 in 2-D is not stored.  No ghost
// cells are stored: the loader maps an out-of-range (column, row) to the
// periodic image or to the boundary condition applied to the owner cell, which
// is what the reference's Halo and Boundary phases leave in the first ghost
// layer (crates/cfd_boundary/src/lib.rs:574-601).
//
// Work decomposition (wave strips): one 64-lane wavefront owns a strip of 64
// consecutive columns and `strip_h` rows.  Lanes 2..61 produce output; lanes
// 0,1,62,63 are halo lanes.  The wave marches up its strip one row at a time:
// x-neighbours come from lane shuffles, y-neighbours from a rolling register
// window, so no LDS and no block barriers are needed.  Each face flux is
// evaluated once per strip (x-faces by the lane right of the face, y-faces by
// the lane of the column), rows are loaded once per strip with 4 halo rows.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "efvm_physics.h"

namespace efvm {

constexpr int WAVE = 64;
constexpr int STRIP_OUT = WAVE - 4;   // output columns per strip
constexpr int STRIP_OUT2 = 2 * WAVE - 4;   // output columns per strip of the two-column periodic path
constexpr int STRIP_OUT23 = 2 * WAVE - 8;  // output columns per strip of the fused stage-2+3 kernel

struct DevClock {
    double lam[2];        // CFL spectral-radius maxima (ping-pong by step parity), as doubles
    double dt;            // dt of the current step
    double time;          // simulated time
    long long step;       // completed steps
    long long bad;        // count of non-finite / non-positive-density cells seen by the CFL scan
    double lam_tmp;       // scratch slot for on-demand CFL queries
};

struct StageArgs {
    int nx, ny;
    int per_x, per_y, seam_mode;
    int bc_kind[4];              // XLo, XHi, YLo, YHi
    double bc_par[4][5];
    double bc_cons[4][5];        // precomputed supersonic-inflow states
    const double* bc_tab[4];     // BC_TABLE: per boundary face (row of an x side, column of a y
                                 // side) rho, rho_u, rho_v, rho_w, rho_E, given (0/1); null: none given
    int parabolic;               // x-lo parabolic velocity inlet (cylinder case)
    double inflow_umax, inflow_h, inflow_rho;
    double lo_x, lo_y, dx, dy, dz;
    double inv_x, inv_y, inv_z;  // face area / cell volume per axis
    double gamma, r_gas;
    int flux, muscl, limiter;
    int viscous, visc_const;
    double mu, dxi2;
    const double* yw;            // per-row viscous weights: yim[ny], yip[ny], yis[ny]
    int body_force;
    double gx;
    const uint8_t* solid;        // nullptr when no immersed body
    int stage;                   // 0, 1, 2 (Euler uses 0 with euler=1)
    int euler;
    int parity;                  // step parity: lam[parity] -> dt, lam[parity^1] <- fused CFL
    const void* uin[4];
    const void* u0[4];
    void* uout[4];
    double cfl, fixed_dt;
    DevClock* clock;
    int fuse_cfl;
    int strip_h, n_strips_x, n_strips_y;
    int xcd_remap;
    double* lam_part;            // FP32 fused CFL: per-wave maxima (null: atomics on the clock)
    unsigned* face_count;        // learned bypass: deferred-face count, cleared by the stage kernel
    unsigned plane_stride;       // bytes from one plane of a set to the next when the four planes of
                                 // every set lie at a common stride (0: they do not)
    int reverse;                 // deal the strips in reverse order (alternate launches sweep the
                                 // grid in opposite directions; results do not depend on the order)
    int pair_path;               // FP32: a two-column strip kernel runs this stage (1 the periodic
                                 // form, 2 the general one; n_strips_x then counts strips of
                                 // STRIP_OUT2 columns)
};

// Which two-column strip kernel applies (0: none, 1: the periodic form, 2: the general form).
// Both need no boundary features (viscous terms, immersed body, parabolic inflow, body force),
// only boundaries whose ghost is a copy, a reflection or a periodic image (outflow, slip, no-slip,
// symmetry, periodic: the EOS-based conditions and ghost tables stay with the one-column
// kernel), and every plane set addressable by one buffer descriptor of under 2 GiB.  The
// periodic form needs both axes periodic, nx even (column pairs never straddle the seam) and
// at least four cells per axis (one wrap per row or column index).  The general form takes any
// mix of those boundaries, with nx even on a periodic x axis, nx >= 2, and ny >= 4 on a
// periodic y axis.
inline int pair_path_kind(int nx, int ny, bool per_x, bool per_y, bool no_features, const int* bc_kind,
                          unsigned plane_stride) {
    if (!no_features || plane_stride == 0 || (unsigned long long)plane_stride * 4ull >= 0x80000000ull) return 0;
    for (int s = 0; s < 4; s++)
        if (bc_kind[s] != BC_OUTFLOW && bc_kind[s] != BC_PERIODIC && bc_kind[s] != BC_SLIP && bc_kind[s] != BC_NOSLIP &&
            bc_kind[s] != BC_SYMMETRY)
            return 0;
    if (per_x && per_y && nx % 2 == 0 && nx >= 4 && ny >= 4) return 1;
    if ((per_x && nx % 2 != 0) || nx < 2 || (per_y && ny < 4)) return 0;
    return 2;
}

// One case of a batched ensemble: its stage arguments and plane sets A, B, C.
struct BatchCase {
    StageArgs a;
    void* P[3][4];
    StageArgs a23;     // the case's arguments in the fused stage-2+3 strip geometry (when fused)
    int fused;         // 1: whole steps run stage 1, then stages 2 and 3 in one fused launch; U^n
                       // then lives in set A at even step parity and in set C at odd (set B holds U1)
};

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <class T> __device__ __forceinline__ T shfl_up1(T v) { return __shfl_up(v, 1, WAVE); }
template <class T> __device__ __forceinline__ T shfl_dn1(T v) { return __shfl_down(v, 1, WAVE); }

// Wave index with an XCD-aware remap: the hardware deals consecutive workgroups
// round-robin over the 8 XCDs; remapping gives each XCD a contiguous range of
// strips, so strips sharing halo rows/columns share one L2.
__device__ __forceinline__ int block_index(const StageArgs& a) {
    int nb = gridDim.x;
    int b = blockIdx.x;
    if (a.xcd_remap && nb >= 8) {
        int per = nb / 8, rem = nb % 8;
        int x = b % 8, k = b / 8;                 // k-th workgroup dealt to XCD x
        if (a.reverse) k = per + (x < rem ? 1 : 0) - 1 - k;
        b = x * per + (x < rem ? x : rem) + k;    // contiguous range per XCD (bijective)
    } else if (a.reverse) {
        b = nb - 1 - b;
    }
    return b;
}
__device__ __forceinline__ int strip_index(const StageArgs& a, int waves_per_block) {
    // wave-uniform: keeps the strip's rows, and every per-row decision, in scalar registers
    return __builtin_amdgcn_readfirstlane(block_index(a) * waves_per_block + (int)(threadIdx.x / WAVE));
}

// Plane element at a wave-uniform base plus a per-lane byte offset (32-bit): the
// load/store takes the base in scalar registers and the offset in one VGPR, so
// the four planes of a row cost no per-lane 64-bit address arithmetic.  The
// row_ld / row_st forms go through a buffer descriptor of one row (`nbytes`),
// built in scalar registers: no vector address arithmetic at all, and an offset
// past the row reads 0 instead of touching memory.
constexpr int BUF_DWORD3 = 0x00020000;   // gfx9 raw buffer: 32-bit data format, no swizzle
__device__ __forceinline__ float row_ld(const float* row, unsigned nbytes, unsigned byte_off) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)row, 0, (int)nbytes, BUF_DWORD3);
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}
__device__ __forceinline__ void row_st(float* row, unsigned nbytes, unsigned byte_off, float v) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)row, 0, (int)nbytes, BUF_DWORD3);
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, byte_off, 0, 0);
}
template <class T>
__device__ __forceinline__ T ld_at(const T* base, unsigned byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ void st_at(T* base, unsigned byte_off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// ---------------------------------------------------------------------------
// loader: interior value, periodic image or boundary-condition ghost
// ---------------------------------------------------------------------------
template <class T> struct Cell { T r, ru, rv, re; };

template <class T>
__device__ __forceinline__ Cell<T> load_plain(const StageArgs& a, const T* const* U, int c, int r) {
    long q = (long)r * a.nx + c;
    return {U[0][q], U[1][q], U[2][q], U[3][q]};
}

// Ghost state for boundary side `side` from its owner cell (cfd_boundary:173-362),
// evaluated in double with the reference's operation order, then rounded to T.
template <class T>
__device__ __forceinline__ Cell<T> bc_apply(const StageArgs& a, int side, Cell<T> in, int owner_row, int owner_col) {
    const int kind = a.bc_kind[side];
    const int axis = side >> 1;
    switch (kind) {
    case BC_TABLE: {
        // FromOutside (crates/cfd_boundary/src/lib.rs:108-147): the state another code supplied
        // for this face, else its fallback SupersonicOutflow (a copy of the owner cell)
        const double* t = a.bc_tab[side];
        if (t) {
            t += 6 * (axis == 0 ? owner_row : owner_col);
            if (t[5] != 0.0) return {T(t[0]), T(t[1]), T(t[2]), T(t[4])};
        }
        return in;
    }
    case BC_SLIP: case BC_SYMMETRY:
        if (axis == 0) in.ru = -in.ru; else in.rv = -in.rv;
        return in;
    case BC_NOSLIP:
        in.ru = -in.ru; in.rv = -in.rv;
        return in;
    case BC_SUP_INFLOW:
        return {T(a.bc_cons[side][0]), T(a.bc_cons[side][1]), T(a.bc_cons[side][2]), T(a.bc_cons[side][4])};
    case BC_SUB_INFLOW: {
        C5<double> u = {double(in.r), double(in.ru), double(in.rv), 0.0, double(in.re)};
        double pin = ref::pressure(u, a.gamma);
        P6<double> p = {a.bc_par[side][0], a.bc_par[side][1], a.bc_par[side][2], a.bc_par[side][3], pin, a.bc_par[side][4]};
        C5<double> o = ref::prim_to_cons(p, a.gamma);
        return {T(o.r), T(o.ru), T(o.rv), T(o.re)};
    }
    case BC_PRESSURE_OUTLET: {
        C5<double> u = {double(in.r), double(in.ru), double(in.rv), 0.0, double(in.re)};
        P6<double> p = ref::cons_to_prim(u, a.gamma, a.r_gas);
        p.p = a.bc_par[side][0];
        p.t = a.bc_par[side][0] / (p.r * 287.058);
        C5<double> o = ref::prim_to_cons(p, a.gamma);
        return {T(o.r), T(o.ru), T(o.rv), T(o.re)};
    }
    default:
        return in;
    }
}

// Ghost state across boundary side `side` from its owner cell: the cylinder case's parabolic
// inflow on x-lo, else the side's stock condition.
template <class T>
__device__ __forceinline__ Cell<T> ghost_of_owner(const StageArgs& a, int side, Cell<T> own, int orow, int oc) {
    if (side == 0 && a.parabolic) {
        // examples/cylinder_drag_validation/main.rs:151-177 (overrides the x-lo fill)
        double y = a.lo_y + ((double)orow + 0.5) * a.dy;
        double um = a.inflow_umax, h = a.inflow_h, rho = a.inflow_rho;
        double uy = 4.0 * um * y * (h - y) / (h * h);
        C5<double> u = {double(own.r), double(own.ru), double(own.rv), 0.0, double(own.re)};
        double pin = ref::pressure(u, a.gamma);
        P6<double> p = {rho, uy, 0.0, 0.0, pin, pin / (rho * 287.058)};
        C5<double> o = ref::prim_to_cons(p, a.gamma);
        return {T(o.r), T(o.ru), T(o.rv), T(o.re)};
    }
    return bc_apply(a, side, own, orow, oc);
}

// Where a cell index outside the grid takes its value from: the periodic image (side -1), or
// the owner cell of the boundary ghost on side 0..3 (x-lo, x-hi, y-lo, y-hi).
struct GhostOwner { int side, col, row; };
__device__ __forceinline__ GhostOwner ghost_owner(const StageArgs& a, int c, int r) {
    const int nx = a.nx, ny = a.ny;
    if (a.per_x) { c %= nx; if (c < 0) c += nx; }
    if (a.per_y) { r %= ny; if (r < 0) r += ny; }
    bool cin = (c >= 0 && c < nx), rin = (r >= 0 && r < ny);
    if (cin && rin) return {-1, c, r};
    if (!cin) return {c < 0 ? 0 : 1, c < 0 ? 0 : nx - 1, r < 0 ? 0 : (r >= ny ? ny - 1 : r)};
    return {r < 0 ? 2 : 3, c, r < 0 ? 0 : ny - 1};
}

// The value at (column c, row r), in the grid or not: the periodic image, or the ghost of the
// owner cell across the boundary (the stage kernels' loaders; kept apart from ghost_owner so
// that the hot paths keep their registers, the combined form spilled in the FP32 kernel).
template <class T>
__device__ __forceinline__ Cell<T> load_cell(const StageArgs& a, const T* const* U, int c, int r) {
    const int nx = a.nx, ny = a.ny;
    if (a.per_x) { c %= nx; if (c < 0) c += nx; }
    if (a.per_y) { r %= ny; if (r < 0) r += ny; }
    bool cin = (c >= 0 && c < nx), rin = (r >= 0 && r < ny);
    if (cin && rin) return load_plain(a, U, c, r);
    if (!cin) {
        int side = c < 0 ? 0 : 1;
        int oc = c < 0 ? 0 : nx - 1;
        int orow = r < 0 ? 0 : (r >= ny ? ny - 1 : r);
        Cell<T> own = load_plain(a, U, oc, orow);
        return ghost_of_owner(a, side, own, orow, oc);
    }
    int side = r < 0 ? 2 : 3;
    int orow = r < 0 ? 0 : ny - 1;
    Cell<T> own = load_plain(a, U, c, orow);
    return bc_apply(a, side, own, orow, c);
}

__device__ __forceinline__ void atomic_max_pos_double(double* addr, double v) {
    // non-negative doubles order like their bit patterns.  The maximum only grows during a
    // launch, so a value not above the current one cannot change it: skipping those keeps
    // thousands of waves that finish together from serialising on one L2 atomic.
    unsigned long long* p = reinterpret_cast<unsigned long long*>(addr);
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    if (b <= __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    atomicMax(p, b);
}

// Wave-wide max of a double via shuffles (xor butterfly).
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// CFL spectral radius of one cell, crates/cfd_solver/src/lib.rs:653-666 (nz = 1:
// w = 0, so the z term is (0 + a)/dz), in the reference's operation order.
__device__ __forceinline__ double cfl_lambda(const StageArgs& a, double r, double ru, double rv, double re) {
    C5<double> u = {r, ru, rv, 0.0, re};
    double v[3]; ref::velocity(u, v);
    double s = ref::sound(u, a.gamma);
    double lam = (fabs(v[0]) + s) / a.dx + (fabs(v[1]) + s) / a.dy + (fabs(v[2]) + s) / a.dz;
    if (a.viscous) {
        P6<double> p = ref::cons_to_prim(u, a.gamma, a.r_gas);
        double mu = a.visc_const ? a.mu : ref::sutherland(p.t);
        double nu = mu / r;
        lam += 2.0 * nu * (1.0 / (a.dx * a.dx) + 1.0 / (a.dy * a.dy) + 1.0 / (a.dz * a.dz));
    }
    return lam;
}

__device__ __forceinline__ double stage_dt(const StageArgs& a) {
    if (a.fixed_dt > 0.0) return a.fixed_dt;
    return a.cfl / a.clock->lam[a.parity];
}

}  // namespace efvm
