// efvm_api.cpp -- implementation of the C ABI declared in include/efvm.h.
//
// A context owns three plane sets (4 SoA planes each) in HBM, in three roles: `cur`
// holds U^n, `prev` U^{n-1}, the third is free.  An RK3 step is
//   stage 1  cur -> prev ;  stage 2  (prev, U0=cur) -> free ;  stage 3  (free, U0=cur) -> prev
// and then cur and prev swap: U^n stays intact through the step (the reference's u0 after
// the step, crates/cfd_solver/src/integrate.rs:76-78), U0 never needs a copy, and the whole
// step is 128 B/cell of HBM traffic in FP32 (256 B in FP64).  Euler ping-pongs cur <-> prev.  dt lives in device
// memory (DevClock): the stage kernels read lambda_max and form dt = cfl/lambda
// themselves, so a step needs no host round trip and can be graph-captured.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/efvm.h"
#include "efvm_stage.h"
#include "efvm_mlp.h"

namespace efvm {
hipError_t launch_stage_f64(const StageArgs& a, int block, hipStream_t st);
hipError_t launch_stage_f32(const StageArgs& a, int block, hipStream_t st);
hipError_t launch_stage_f32_pair23(const StageArgs& a, hipStream_t st);
template <class T> hipError_t launch_cfl(const StageArgs& a, const void* const* planes, double* target, hipStream_t st);
hipError_t launch_clock_set(DevClock* clk, double dt, double time, long long step, hipStream_t st);
struct IbmArgs {
    int nx, ny;
    double lo_x, lo_y, dx, dy, dz;
    double gamma, r_gas;
    int per_x, per_y;
    const uint8_t* solid;
};
template <class T> hipError_t launch_ibm_mirror(const StageArgs& sa, const IbmArgs& a, int n, const int* idx, const double* img,
                                                void* pre, hipStream_t st);
template <class T> hipError_t launch_wall_force(const StageArgs& sa, const IbmArgs& a, int n, const int* idx, int n_src,
                                                void* const* set0, void* const* set1, const void* pre0, const void* pre1,
                                                double w0, double w1, int visc_const, double mu, double* out, hipStream_t st);
template <class T> hipError_t launch_marker_force(const StageArgs& sa, const IbmArgs& a, int n, const double* mk, double off,
                                                  int visc_const, double mu, double* out, hipStream_t st);
template <class T> hipError_t launch_sponge(const StageArgs& a, double x0, double x1, double sig, double um, double h,
                                            double rho, double p, hipStream_t st);
template <class T> hipError_t launch_sums(const StageArgs& a, double* part, int nb, double vol, double* out, hipStream_t st);
template <class T> hipError_t launch_sums_partial(const StageArgs& a, double* part, int nb, hipStream_t st);
hipError_t launch_ibm_flip(void* const* planes, int esz, const uint8_t* solid, long n, const double* stagnant, hipStream_t st);
hipError_t launch_stage_f32_faces(const StageArgs& a, const float* F, hipStream_t st);
hipError_t launch_stage_f32_flux(const StageArgs& a, const float* F, hipStream_t st);
hipError_t launch_lam_reduce(const double* part, int n, double* target, hipStream_t st);
hipError_t launch_batch_lam_fold(const BatchCase* cases, int n_cases, int parity, hipStream_t st);
hipError_t launch_cfl_f32(const StageArgs& a, const void* const* planes, double* target, hipStream_t st);
hipError_t launch_mlp_faces(const mlp::FaceArgs& a, hipStream_t st);
hipError_t launch_mlp_faces_bypass(const mlp::FaceArgs& a, hipStream_t st);
hipError_t launch_mlp_light_stage(const StageArgs& a, const mlp::FaceArgs& fa, int blocks, hipStream_t st);
hipError_t launch_mlp_infer(const mlp::Packed* P, const float* X, long n, float* Y, hipStream_t st);
namespace mlp { struct TrainArgs; struct SampleArgs; }
hipError_t launch_mlp_train(const mlp::TrainArgs& a, int nblk, float* grad, double* loss_out, hipStream_t st);
hipError_t launch_mlp_adam(float* p, const float* g, float* m, float* v, float lr, int t, int n_update, hipStream_t st);
hipError_t launch_mlp_samples(const mlp::SampleArgs& a, long n_out, hipStream_t st);
hipError_t launch_mlp_gram(const float* X, const float* Y, long n, long stride, double* partial, int blocks,
                           hipStream_t st);
}  // namespace efvm

using namespace efvm;

struct efvm_ctx {
    efvm_grid grid{};
    efvm_gas gas{};
    efvm_scheme sch{};
    efvm_bc bc[4]{};
    int device = 0;
    hipStream_t stream = nullptr;
    size_t esz = 8;
    long ncell = 0;
    void* P[3][4] = {};            // three plane sets of four planes
    void* set_base[3] = {};        // each set is one allocation: plane m at set_base + m * plane_stride
    size_t plane_stride = 0;
    int cur = 0;                   // index of the set holding U^n
    int prev = 1;                  // index of the set holding U^{n-1} (U1 / U2's input during a step)
    int have_prev = 0;             // prev holds the state before the last completed step (or a loaded u0)
    int ghost_hist = 0;            // the last completed step's substage inputs and pre-mirror ghosts are resident
    DevClock* clock = nullptr;
    double* lam_part = nullptr;    // FP32: per-wave maxima of the fused CFL
    long lam_slots = 0;            // entries of lam_part
    double* bc_tab[4] = {nullptr, nullptr, nullptr, nullptr};   // EFVM_BC_TABLE ghost tables
    double* yw = nullptr;
    uint8_t* solid = nullptr;
    int* ghost_idx = nullptr;
    double* ghost_img = nullptr;
    void* ghost_pre[2] = {};       // each ghost before the IBM mirror: of the first / the latest later substage
    int n_ghost = 0;
    double* markers = nullptr;
    int n_mark = 0;
    double* scratch = nullptr;     // sums partials (1024*4) + 6 marker + 5 sums
    double* stagnant = nullptr;
    int ibm_flipped = 0;
    int rk_stage = 0;
    int parity = 0;
    int lam_valid = 0;
    int strip_h = 31;   // rows per strip; the FP32 kernel runs row steps in groups of four (strip_h + 1 of them)
    int waves_per_block = 4;
    int allow_pair = 1;            // FP32 periodic grids may use the two-column strip kernel
    int allow_fuse23 = 1;          // ... and run RK3 stages 2 and 3 as one fused launch (seam mode 1)
    int strip_h23 = 73;            // rows per strip of the fused stage-2+3 kernel
    int xcd_remap = 1;
    // consecutive stage launches sweep the strips in opposite directions, so each launch starts
    // where the previous one ended, on rows whose lines the device's last-level cache still holds
    unsigned dir_count = 0;
    int parabolic = 0;
    double inflow[3] = {0, 0, 0};
    int sponge = 0;
    double sp[7] = {0, 0, 0, 0, 0, 0, 0};
    int body_force = 0;
    double gx = 0.0;
    int n_cyl = 0;
    double cyl[8][3] = {};
    double freestream[5] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
    int graph_cur = 0, graph_parity = 0;   // plane set and CFL slot the captured steps start from
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int timed = 0;
    void* snap = nullptr;          // autotuner snapshot (state + clock)
    int snap_parity = 0, snap_lam = 0, snap_flip = 0, snap_cur = 0, snap_prev = 1, snap_have_prev = 0;
    int learned = 0;               // learned face reconstruction attached
    int learned_bypass = 0;        // gate evaluated only where the ungated face value is near the TVD bound
    int learned_own4 = 0;          // the attached network's correction is the 4-tap own-variable form
    float learned_taps[4][4] = {}; // its open-gate taps
    float* faces = nullptr;        // 16 planes of face primitives
    unsigned* dlist = nullptr;     // bypass: deferred faces (2 nx ny entries), their count, then the fused
                                   // stage's pending-cell count and finished-workgroup counter
    int learned_fused = 1;         // bypass stages on the 4-tap correction run as one fused stage kernel
    mlp::Packed* wpack = nullptr;  // the attached network's inference weights
    unsigned long long* lstats = nullptr;   // [fallback faces, faces evaluated, sides the network gated]
    std::string err;
};

namespace {

efvm_status fail(efvm_ctx* c, efvm_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}
#define HIPCHK(ctx, expr)                                                                              \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) return fail(ctx, EFVM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

bool periodic_axis(const efvm_ctx* c, int axis) { return c->bc[2 * axis].kind == EFVM_BC_PERIODIC; }

// FP32 periodic grids without boundary features can run the two-column strip kernel.
int pair_fits(const efvm_ctx* c, const StageArgs& a) {
    const bool plain = !a.viscous && !a.solid && !a.parabolic && !a.body_force;
    if (c->esz != 4 || c->learned) return 0;
    return pair_path_kind(a.nx, a.ny, a.per_x, a.per_y, plain, a.bc_kind, a.plane_stride);
}

StageArgs base_args(const efvm_ctx* c) {
    StageArgs a;
    std::memset(&a, 0, sizeof(a));
    const efvm_grid& g = c->grid;
    a.nx = g.nx; a.ny = g.ny;
    a.per_x = periodic_axis(c, 0); a.per_y = periodic_axis(c, 1);
    a.seam_mode = c->sch.seam_mode;
    for (int s = 0; s < 4; s++) {
        a.bc_kind[s] = c->bc[s].kind;
        a.bc_tab[s] = c->bc[s].kind == EFVM_BC_TABLE ? c->bc_tab[s] : nullptr;
        for (int q = 0; q < 5; q++) a.bc_par[s][q] = c->bc[s].par[q];
        if (c->bc[s].kind == EFVM_BC_SUPERSONIC_INFLOW) {
            const double* p = c->bc[s].par;
            P6<double> pv = {p[0], p[1], p[2], p[3], p[4], p[4] / (p[0] * 287.058)};
            C5<double> o = ref::prim_to_cons(pv, c->gas.gamma);
            a.bc_cons[s][0] = o.r; a.bc_cons[s][1] = o.ru; a.bc_cons[s][2] = o.rv; a.bc_cons[s][3] = o.rw; a.bc_cons[s][4] = o.re;
        }
    }
    a.parabolic = c->parabolic;
    a.inflow_umax = c->inflow[0]; a.inflow_h = c->inflow[1]; a.inflow_rho = c->inflow[2];
    a.lo_x = g.lo_x; a.lo_y = g.lo_y;
    a.dx = (g.hi_x - g.lo_x) / (double)g.nx;
    a.dy = (g.hi_y - g.lo_y) / (double)g.ny;
    a.dz = g.dz;
    double vol = a.dx * a.dy * a.dz;
    a.inv_x = (a.dy * a.dz) / vol;
    a.inv_y = (a.dx * a.dz) / vol;
    a.inv_z = (a.dx * a.dy) / vol;
    a.gamma = c->gas.gamma; a.r_gas = c->gas.r_gas;
    a.flux = c->sch.flux; a.muscl = c->sch.recon == EFVM_RECON_MUSCL; a.limiter = c->sch.limiter;
    a.viscous = c->sch.viscous; a.visc_const = c->gas.viscosity == EFVM_VISC_CONSTANT; a.mu = c->gas.mu;
    a.dxi2 = 1.0 / (a.dx * a.dx);
    a.yw = c->yw;
    a.body_force = c->body_force; a.gx = c->gx;
    a.solid = c->n_cyl > 0 ? c->solid : nullptr;
    a.euler = c->sch.integrator == EFVM_INTEGRATOR_EULER;
    a.cfl = c->sch.cfl; a.fixed_dt = c->sch.fixed_dt;
    a.clock = c->clock;
    a.strip_h = c->strip_h;
    a.n_strips_x = (g.nx + STRIP_OUT - 1) / STRIP_OUT;
    a.n_strips_y = (g.ny + c->strip_h - 1) / c->strip_h;
    a.xcd_remap = c->xcd_remap;
    a.lam_part = c->lam_part;
    a.face_count = nullptr;
    a.plane_stride = c->plane_stride <= 0xFFFFFFFFu ? (unsigned)c->plane_stride : 0u;
    a.pair_path = c->allow_pair ? pair_fits(c, a) : 0;
    if (a.pair_path) a.n_strips_x = (g.nx + STRIP_OUT2 - 1) / STRIP_OUT2;
    return a;
}

IbmArgs ibm_args(const efvm_ctx* c, const StageArgs& a) {
    IbmArgs b;
    b.nx = a.nx; b.ny = a.ny; b.lo_x = a.lo_x; b.lo_y = a.lo_y; b.dx = a.dx; b.dy = a.dy; b.dz = a.dz;
    b.gamma = a.gamma; b.r_gas = a.r_gas; b.per_x = a.per_x; b.per_y = a.per_y; b.solid = c->solid;
    return b;
}

// The plane set holding the state of the current substage: U^n at a step boundary, U1 / U2
// after the first / second RK3 substage.
int stage_set(const efvm_ctx* c) {
    if (c->sch.integrator == EFVM_INTEGRATOR_RK3 && c->rk_stage == 1) return c->prev;
    if (c->sch.integrator == EFVM_INTEGRATOR_RK3 && c->rk_stage == 2) return 3 - c->cur - c->prev;
    return c->cur;
}

void set_planes(StageArgs& a, void* const* in, void* const* u0, void* const* out) {
    for (int m = 0; m < 4; m++) { a.uin[m] = in[m]; a.u0[m] = u0[m]; a.uout[m] = out[m]; }
}

efvm_status launch_stage(efvm_ctx* c, const StageArgs& a) {
    int block = WAVE * c->waves_per_block;
    hipError_t e = c->esz == 8 ? launch_stage_f64(a, block, c->stream) : launch_stage_f32(a, block, c->stream);
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("stage kernel launch: ") + hipGetErrorString(e));
    return EFVM_OK;
}

efvm_status launch_cfl_into(efvm_ctx* c, void* const* planes, double* target) {
    StageArgs a = base_args(c);
    // FP32: the fused epilogues' per-cell formula, so a restarted run takes the same dt
    hipError_t e = c->esz == 8 ? launch_cfl<double>(a, planes, target, c->stream) : launch_cfl_f32(a, planes, target, c->stream);
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("cfl kernel: ") + hipGetErrorString(e));
    return EFVM_OK;
}

// After the last substage of a step: U^{n+1} is in the old prev set and U^n becomes prev; the
// sponge, the CFL of the new state (unless the stage fused it), the CFL slot parity.  u2_kept:
// U2 is in the third set (not after a fused step), which the wall force's ghost cells read.
efvm_status end_of_step(efvm_ctx* c, bool cfl_fused, int u2_kept) {
    std::swap(c->cur, c->prev);
    c->have_prev = 1;
    c->ghost_hist = u2_kept;
    StageArgs b = base_args(c);
    set_planes(b, c->P[c->cur], c->P[c->cur], c->P[c->cur]);
    if (c->sponge) {
        hipError_t e = c->esz == 8 ? launch_sponge<double>(b, c->sp[0], c->sp[1], c->sp[2], c->sp[3], c->sp[4], c->sp[5], c->sp[6], c->stream)
                                   : launch_sponge<float>(b, c->sp[0], c->sp[1], c->sp[2], c->sp[3], c->sp[4], c->sp[5], c->sp[6], c->stream);
        if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, "sponge");
    }
    if (c->sch.fixed_dt <= 0.0 && !cfl_fused) {
        efvm_status st = launch_cfl_into(c, c->P[c->cur], &c->clock->lam[c->parity ^ 1]);
        if (st != EFVM_OK) return st;
    }
    c->parity ^= 1;
    c->rk_stage = 0;
    return EFVM_OK;
}

// One substage (one app.run()).  Enqueue only; no host synchronisation.
// Bypass stages whose correction is the 4-tap form run as one fused stage kernel and the
// pending-cell pass (cells are listed as index << 4 | face mask: at most 2^28 cells).
bool learned_fused_fits(const efvm_ctx* c) {
    return c->learned && c->learned_bypass && c->learned_fused && c->learned_own4 && c->dlist && c->lam_part &&
           c->ncell < (1L << 28);
}

efvm_status run_substage(efvm_ctx* c) {
    const bool rk3 = c->sch.integrator == EFVM_INTEGRATOR_RK3;
    // RK3: A = U^n (set cur, never rotates), B = U1, C = U2.  Euler ping-pongs sets 0 <-> 1,
    // so an even number of steps returns the state to its starting set (graph replay).
    void* const* A = c->P[c->cur];
    void* const* B = c->P[c->prev];
    void* const* Cc = c->P[3 - c->cur - c->prev];
    efvm_status st;
    // Setup: CFL of U^n (first substage only) ...
    if (c->rk_stage == 0 && !c->lam_valid && c->sch.fixed_dt <= 0.0) {
        if ((st = launch_cfl_into(c, A, &c->clock->lam[c->parity])) != EFVM_OK) return st;
        c->lam_valid = 1;
    }
    // ... then the cylinder mask refresh: cells inside a body become the stagnant state (once, static body).
    if (c->n_cyl > 0 && !c->ibm_flipped) {
        hipError_t e = launch_ibm_flip(A, (int)c->esz, c->solid, c->ncell, c->stagnant, c->stream);
        if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, "ibm flip");
        c->ibm_flipped = 1;
    }
    StageArgs a = base_args(c);
    a.parity = c->parity;
    a.stage = rk3 ? c->rk_stage : 0;
    a.reverse = (int)(c->dir_count++ & 1u);
    void* const* in = A;
    void* const* out = B;
    if (rk3 && c->rk_stage == 1) { in = B; out = Cc; }
    if (rk3 && c->rk_stage == 2) { in = Cc; out = B; }
    set_planes(a, in, A, out);
    const bool last = !rk3 || c->rk_stage == 2;
    a.fuse_cfl = last && rk3 && !c->sponge && c->sch.fixed_dt <= 0.0;
    // PreFlux: IBM mirror into the stage input
    if (c->n_cyl > 0 && c->n_ghost > 0) {
        IbmArgs ib = ibm_args(c, a);
        hipError_t e = c->esz == 8 ? launch_ibm_mirror<double>(a, ib, c->n_ghost, c->ghost_idx, c->ghost_img, c->ghost_pre[c->rk_stage ? 1 : 0], c->stream)
                                   : launch_ibm_mirror<float>(a, ib, c->n_ghost, c->ghost_idx, c->ghost_img, c->ghost_pre[c->rk_stage ? 1 : 0], c->stream);
        if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, "ibm mirror");
    }
    bool fused_light = false;   // the fused learned stage folds its CFL slots itself
    if (c->learned) {
        // face states of the stage input from the network (with the MUSCL fallback), then the stage
        mlp::FaceArgs fa;
        fa.nx = c->grid.nx; fa.ny = c->grid.ny;
        fa.n_chunks = (c->grid.nx + 63) / 64;
        fa.n_tasks = 2 * c->grid.ny * fa.n_chunks;
        fa.gamma = (float)c->gas.gamma;
        for (int m = 0; m < 4; m++) fa.U[m] = (const float*)in[m];
        fa.F = c->faces;
        fa.w = c->wpack;
        fa.stats = c->lstats;
        fa.bypass = c->learned_bypass;
        // the bypass passes finish each face with its flux (they run at full occupancy); the
        // network pass writes face states and the stage kernel evaluates the fluxes
        fa.out_mode = c->learned_bypass ? 1 : 0;
        fa.flux = c->sch.flux;
        fa.list = c->dlist;
        fa.n_list = c->dlist ? c->dlist + 2 * c->ncell : nullptr;
        fa.own4 = c->learned_own4;
        fa.n_lam_slots = 0;
        std::memcpy(fa.taps, c->learned_taps, sizeof(fa.taps));
        if (learned_fused_fits(c)) {
            // one fused stage kernel (faces, fluxes, update) and the pending-cell pass
            StageArgs b = a;
            b.strip_h = mlp::LIGHT_ROWS;
            b.n_strips_x = (c->grid.nx + mlp::LIGHT_OUT - 1) / mlp::LIGHT_OUT;
            b.n_strips_y = (c->grid.ny + mlp::LIGHT_ROWS - 1) / mlp::LIGHT_ROWS;
            fa.n_list = c->dlist + 2 * c->ncell + 1;
            const long tiles = (long)b.n_strips_x * b.n_strips_y;
            const int blocks = (int)std::min<long>(tiles, std::min<long>(mlp::STAT_SLOTS, c->lam_slots) / (mlp::LIGHT_ROWS + 1));
            // the pending-cell pass folds the stage's per-wave CFL maxima (no reduction launch)
            fa.n_lam_slots = (mlp::LIGHT_ROWS + 1) * blocks;
            fused_light = true;
            hipError_t e = launch_mlp_light_stage(b, fa, blocks, c->stream);
            if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("fused learned stage: ") + hipGetErrorString(e));
        } else {
            hipError_t e = c->learned_bypass ? launch_mlp_faces_bypass(fa, c->stream) : launch_mlp_faces(fa, c->stream);
            if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("face kernels: ") + hipGetErrorString(e));
            a.face_count = c->learned_bypass ? fa.n_list : nullptr;
            e = fa.out_mode == 1 ? launch_stage_f32_flux(a, c->faces, c->stream) : launch_stage_f32_faces(a, c->faces, c->stream);
            if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("learned stage kernel: ") + hipGetErrorString(e));
        }
    } else if ((st = launch_stage(c, a)) != EFVM_OK) {
        return st;
    }
    if (a.fuse_cfl && a.lam_part && !fused_light) {
        // the stage left one CFL maximum per wave: strips, or 64 cells of the learned stage
        const int nw = c->learned ? (int)((c->ncell + 63) / 64) : a.n_strips_x * a.n_strips_y;
        hipError_t e = launch_lam_reduce(a.lam_part, nw, &c->clock->lam[a.parity ^ 1], c->stream);
        if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("cfl reduction: ") + hipGetErrorString(e));
    }
    if (last) return end_of_step(c, a.fuse_cfl != 0, 1);
    c->rk_stage++;
    return EFVM_OK;
}

// Whole steps.  FP32 RK3 on the periodic two-column path runs stage 1, then
// stages 2 and 3 fused (stage_f32_pair23_kernel): U1 goes to the third set, U^{n+1} to prev,
// and the roles swap as after three substages; results are bitwise the same.
bool fused_path_fits(const efvm_ctx* c, const StageArgs& a) {
    return c->esz == 4 && !c->learned && c->sch.integrator == EFVM_INTEGRATOR_RK3 && a.pair_path == 1 &&
           a.ny >= 8 && c->n_cyl == 0;
}
bool fused_step_fits(const efvm_ctx* c, const StageArgs& a) {
    return c->allow_fuse23 && c->rk_stage == 0 && fused_path_fits(c, a);
}

StageArgs fused_args(const efvm_ctx* c, StageArgs a) {
    a.strip_h = c->strip_h23;
    a.n_strips_x = (a.nx + STRIP_OUT23 - 1) / STRIP_OUT23;
    a.n_strips_y = (a.ny + c->strip_h23 - 1) / c->strip_h23;
    return a;
}

efvm_status run_step_fused(efvm_ctx* c) {
    void* const* A = c->P[c->cur];
    void* const* B = c->P[c->prev];
    void* const* Cc = c->P[3 - c->cur - c->prev];
    efvm_status st;
    if (!c->lam_valid && c->sch.fixed_dt <= 0.0) {
        if ((st = launch_cfl_into(c, A, &c->clock->lam[c->parity])) != EFVM_OK) return st;
        c->lam_valid = 1;
    }
    StageArgs a = base_args(c);
    a.parity = c->parity;
    a.stage = 0;
    a.reverse = (int)(c->dir_count++ & 1u);
    set_planes(a, A, A, Cc);             // stage 1: U^n -> U1 in the third set
    if ((st = launch_stage(c, a)) != EFVM_OK) return st;
    StageArgs b = fused_args(c, a);
    b.stage = 2;
    b.reverse = (int)(c->dir_count++ & 1u);
    set_planes(b, Cc, A, B);             // stages 2 and 3: U1, U^n -> U^{n+1} in prev
    b.fuse_cfl = !c->sponge && c->sch.fixed_dt <= 0.0;
    // Each workgroup takes its strip's CFL maximum into the clock with an atomic maximum. The
    // launch runs its workgroups over several rounds, so the atomics arrive spread out, and the
    // step needs no slot-fold launch: 4.5 us less per 4096^2 step than the slots. (Thousands of
    // same-address atomics from a one-round launch, whose workgroups finish together, serialise.
    // That is why the substage launches keep their slots.)
    b.lam_part = nullptr;
    hipError_t e = launch_stage_f32_pair23(b, c->stream);
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("fused stage kernel launch: ") + hipGetErrorString(e));
    return end_of_step(c, b.fuse_cfl != 0, 0);
}

efvm_status run_step(efvm_ctx* c) {
    if (fused_step_fits(c, base_args(c))) return run_step_fused(c);
    const int nsub = c->sch.integrator == EFVM_INTEGRATOR_RK3 ? 3 : 1;
    for (int q = 0; q < nsub; q++) {
        efvm_status st = run_substage(c);
        if (st != EFVM_OK) return st;
    }
    return EFVM_OK;
}

efvm_status device_alloc(efvm_ctx* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return EFVM_OK;
}

void free_ibm(efvm_ctx* c) {
    if (c->solid) (void)hipFree(c->solid);
    if (c->ghost_idx) (void)hipFree(c->ghost_idx);
    if (c->ghost_img) (void)hipFree(c->ghost_img);
    for (int k = 0; k < 2; k++) if (c->ghost_pre[k]) (void)hipFree(c->ghost_pre[k]);
    if (c->markers) (void)hipFree(c->markers);
    if (c->stagnant) (void)hipFree(c->stagnant);
    c->solid = nullptr; c->ghost_idx = nullptr; c->ghost_img = nullptr; c->ghost_pre[0] = c->ghost_pre[1] = nullptr; c->markers = nullptr;
    c->stagnant = nullptr;
}

void drop_graph(efvm_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr; c->graph = nullptr; c->graph_steps = 0;
}

}  // namespace

extern "C" {

efvm_status efvm_ctx_create(const efvm_grid* grid, const efvm_gas* gas, const efvm_scheme* scheme, efvm_ctx** out) {
    if (!grid || !gas || !scheme || !out) return EFVM_ERR_ARG;
    *out = nullptr;
    if (grid->nx < 1 || grid->ny < 1 || !(grid->hi_x > grid->lo_x) || !(grid->hi_y > grid->lo_y) || !(grid->dz > 0.0))
        return EFVM_ERR_CONFIG;
    if (scheme->flux < 0 || scheme->flux > 2 || scheme->recon < 0 || scheme->recon > 1 || scheme->limiter < 0 ||
        scheme->limiter > 1 || scheme->integrator < 0 || scheme->integrator > 1 || scheme->precision < 0 || scheme->precision > 1)
        return EFVM_ERR_CONFIG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EFVM_ERR_NO_DEVICE;
    efvm_ctx* c = new (std::nothrow) efvm_ctx();
    if (!c) return EFVM_ERR_ARG;
    c->grid = *grid; c->gas = *gas; c->sch = *scheme;
    for (int s = 0; s < 4; s++) { c->bc[s].kind = EFVM_BC_OUTFLOW; for (int q = 0; q < 5; q++) c->bc[s].par[q] = 0.0; }
    c->esz = scheme->precision == EFVM_FP64 ? 8 : 4;
    c->ncell = (long)grid->nx * grid->ny;
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return EFVM_ERR_HIP;
    }
    efvm_status st = EFVM_OK;
    // each set in one allocation, planes 256-byte aligned, with 4 KiB + 256 B of padding: planes
    // a power of two apart put the four same-cell accesses of a row on the same memory channels
    // (measured 20% slower at 4096^2)
    c->plane_stride = ((c->esz * c->ncell + 255) / 256) * 256 + 4096 + 256;
    for (int s = 0; s < 3 && st == EFVM_OK; s++) {
        st = device_alloc(c, &c->set_base[s], 4 * c->plane_stride);
        for (int m = 0; m < 4 && st == EFVM_OK; m++) c->P[s][m] = (char*)c->set_base[s] + m * c->plane_stride;
    }
    if (st == EFVM_OK) st = device_alloc(c, (void**)&c->clock, sizeof(DevClock));
    if (st == EFVM_OK && c->esz == 4) {
        // slots for the most waves a fused-CFL stage can launch: strips of one row, 64 cells, or
        // the fused learned stage's workgroups of LIGHT_ROWS + 1 waves (at most STAT_SLOTS waves)
        const long strips = (long)((grid->nx + STRIP_OUT - 1) / STRIP_OUT) * grid->ny;
        const long light = std::min<long>((long)((grid->nx + mlp::LIGHT_OUT - 1) / mlp::LIGHT_OUT) *
                                              ((grid->ny + mlp::LIGHT_ROWS - 1) / mlp::LIGHT_ROWS) * (mlp::LIGHT_ROWS + 1),
                                          mlp::STAT_SLOTS);
        const long n = std::max(std::max(strips, light), (c->ncell + 63) / 64) + 64;
        st = device_alloc(c, (void**)&c->lam_part, sizeof(double) * n);
        c->lam_slots = n;
    }
    if (st == EFVM_OK) st = device_alloc(c, (void**)&c->yw, sizeof(double) * 3 * grid->ny);
    if (st == EFVM_OK) st = device_alloc(c, (void**)&c->scratch, sizeof(double) * (1024 * 4 + 16));
    if (st != EFVM_OK) { efvm_ctx_destroy(c); return st; }
    for (int s = 0; s < 3; s++)
        for (int m = 0; m < 4; m++) (void)hipMemsetAsync(c->P[s][m], 0, c->esz * c->ncell, c->stream);
    // viscous y weights from cell-centre distances (crates/cfd_solver/src/lib.rs:483-493)
    std::vector<double> yw(3 * grid->ny);
    double dy = (grid->hi_y - grid->lo_y) / (double)grid->ny;
    auto yc = [&](int j) { return grid->lo_y + ((double)j + 0.5) * dy; };
    for (int j = 0; j < grid->ny; j++) {
        double hm = yc(j) - yc(j - 1), hp = yc(j + 1) - yc(j);
        yw[j] = 1.0 / hm; yw[grid->ny + j] = 1.0 / hp; yw[2 * grid->ny + j] = 2.0 / (hm + hp);
    }
    (void)hipMemcpyAsync(c->yw, yw.data(), sizeof(double) * yw.size(), hipMemcpyHostToDevice, c->stream);
    (void)launch_clock_set(c->clock, 0.0, 0.0, 0, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { efvm_ctx_destroy(c); return EFVM_ERR_HIP; }
    *out = c;
    return EFVM_OK;
}

void efvm_ctx_destroy(efvm_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (int s = 0; s < 3; s++) if (c->set_base[s]) (void)hipFree(c->set_base[s]);
    if (c->clock) (void)hipFree(c->clock);
    if (c->lam_part) (void)hipFree(c->lam_part);
    for (int s = 0; s < 4; s++)
        if (c->bc_tab[s]) (void)hipFree(c->bc_tab[s]);
    if (c->yw) (void)hipFree(c->yw);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->snap) (void)hipFree(c->snap);
    if (c->faces) (void)hipFree(c->faces);
    if (c->wpack) (void)hipFree(c->wpack);
    if (c->lstats) (void)hipFree(c->lstats);
    if (c->dlist) (void)hipFree(c->dlist);
    free_ibm(c);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* efvm_last_error(const efvm_ctx* c) { return c ? c->err.c_str() : "null context"; }

efvm_status efvm_set_launch(efvm_ctx* c, int strip_rows, int waves_per_block) {
    if (!c) return EFVM_ERR_ARG;
    if (strip_rows < 0 || strip_rows > 4096 || waves_per_block < 0 || waves_per_block > 4)
        return fail(c, EFVM_ERR_ARG, "launch shape out of range");
    if (strip_rows) c->strip_h = strip_rows;
    if (waves_per_block) c->waves_per_block = waves_per_block;
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_get_launch(const efvm_ctx* c, int* strip_rows, int* waves_per_block, int* kernel, int* pair_fits_out) {
    if (!c) return EFVM_ERR_ARG;
    if (strip_rows) *strip_rows = c->strip_h;
    if (waves_per_block) *waves_per_block = c->waves_per_block;
    if (kernel) *kernel = c->learned ? 3 : (c->esz == 8 ? 2 : (base_args(c).pair_path ? 1 : 0));
    if (pair_fits_out) *pair_fits_out = pair_fits(c, base_args(c));
    return EFVM_OK;
}

efvm_status efvm_set_fused_step(efvm_ctx* c, int allow, int strip_rows) {
    if (!c) return EFVM_ERR_ARG;
    if (strip_rows < 0 || strip_rows > 4096) return fail(c, EFVM_ERR_ARG, "fused strip height out of range");
    c->allow_fuse23 = allow ? 1 : 0;
    if (strip_rows) c->strip_h23 = strip_rows;
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_get_fused_step(const efvm_ctx* c, int* allow, int* strip_rows, int* fits) {
    if (!c) return EFVM_ERR_ARG;
    if (allow) *allow = c->allow_fuse23;
    if (strip_rows) *strip_rows = c->strip_h23;
    if (fits) *fits = fused_path_fits(c, base_args(c)) ? 1 : 0;
    return EFVM_OK;
}

efvm_status efvm_device_arch(int device, char* arch, size_t n) {
    if (!arch || n == 0) return EFVM_ERR_ARG;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return EFVM_ERR_NO_DEVICE;
    std::snprintf(arch, n, "%s", p.gcnArchName);
    return EFVM_OK;
}

efvm_status efvm_set_pair_kernel(efvm_ctx* c, int allow) {
    if (!c || (allow != 0 && allow != 1)) return EFVM_ERR_ARG;
    c->allow_pair = allow;
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_set_boundaries(efvm_ctx* c, const efvm_bc sides[4]) {
    if (!c || !sides) return EFVM_ERR_ARG;
    for (int s = 0; s < 4; s++)
        if (sides[s].kind < 0 || sides[s].kind > EFVM_BC_TABLE) return fail(c, EFVM_ERR_CONFIG, "unknown boundary kind");
    for (int ax = 0; ax < 2; ax++)
        if ((sides[2 * ax].kind == EFVM_BC_PERIODIC) != (sides[2 * ax + 1].kind == EFVM_BC_PERIODIC))
            return fail(c, EFVM_ERR_CONFIG, std::string("periodic boundary on axis ") + std::to_string(ax) +
                                                " requires both sides periodic (crates/cfd_boundary/src/lib.rs:437-448)");
    for (int s = 0; s < 4; s++) c->bc[s] = sides[s];
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_set_ghost_table(efvm_ctx* c, int side, const double* cons5, const unsigned char* given, size_t n) {
    if (!c || !cons5 || !given || side < 0 || side > 3) return EFVM_ERR_ARG;
    if (c->bc[side].kind != EFVM_BC_TABLE) return fail(c, EFVM_ERR_CONFIG, "ghost table on a side whose kind is not EFVM_BC_TABLE");
    const size_t nf = side < 2 ? (size_t)c->grid.ny : (size_t)c->grid.nx;
    if (n != nf) return fail(c, EFVM_ERR_ARG, "ghost table: one entry per boundary face (ny for x sides, nx for y sides)");
    std::vector<double> h(6 * nf);
    for (size_t k = 0; k < nf; k++) {
        if (given[k] && cons5[5 * k + 3] != 0.0) return fail(c, EFVM_ERR_UNSUPPORTED, "rho_w must be 0 on the 2-D GPU path");
        for (int m = 0; m < 5; m++) h[6 * k + m] = cons5[5 * k + m];
        h[6 * k + 5] = given[k] ? 1.0 : 0.0;
    }
    if (!c->bc_tab[side]) {
        HIPCHK(c, hipMalloc((void**)&c->bc_tab[side], sizeof(double) * 6 * nf));
        drop_graph(c);   // a captured graph holds the stage arguments without this table
    }
    HIPCHK(c, hipMemcpyAsync(c->bc_tab[side], h.data(), sizeof(double) * 6 * nf, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm_set_parabolic_inflow(efvm_ctx* c, double u_max, double height, double rho) {
    if (!c) return EFVM_ERR_ARG;
    c->parabolic = 1; c->inflow[0] = u_max; c->inflow[1] = height; c->inflow[2] = rho;
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_set_sponge(efvm_ctx* c, double x0, double x1, double sigma, double u_max, double height, double rho, double p) {
    if (!c) return EFVM_ERR_ARG;
    if (c->sch.integrator != EFVM_INTEGRATOR_RK3) return fail(c, EFVM_ERR_UNSUPPORTED, "sponge needs RK3");
    c->sponge = 1;
    double v[7] = {x0, x1, sigma, u_max, height, rho, p};
    std::memcpy(c->sp, v, sizeof(v));
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_set_time_step(efvm_ctx* c, double cfl, double fixed_dt) {
    if (!c || !(cfl > 0.0)) return EFVM_ERR_ARG;
    if (c->rk_stage != 0) return fail(c, EFVM_ERR_ARG, "time-step control changes at a step boundary only");
    c->sch.cfl = cfl;
    c->sch.fixed_dt = fixed_dt;
    c->lam_valid = 0;              // back in CFL mode the next step rescans the state
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_set_body_force(efvm_ctx* c, double gx) {
    if (!c) return EFVM_ERR_ARG;
    c->body_force = 1; c->gx = gx;
    drop_graph(c);
    return EFVM_OK;
}

static efvm_status upload_set(efvm_ctx* c, int set, const double* cons5, size_t n);

efvm_status efvm_upload_state(efvm_ctx* c, const double* cons5, size_t n) {
    if (!c || !cons5) return EFVM_ERR_ARG;
    if ((long)n != c->ncell) return fail(c, EFVM_ERR_ARG, "upload: n_cells != nx*ny");
    efvm_status st = upload_set(c, c->cur, cons5, n);
    if (st != EFVM_OK) return st;
    c->rk_stage = 0;
    c->lam_valid = 0;
    c->have_prev = 0;
    c->ghost_hist = 0;
    return EFVM_OK;
}

// Wire-order host state -> the planes of set `set` (synchronous).
static efvm_status upload_set(efvm_ctx* c, int set, const double* cons5, size_t n) {
    std::vector<char> h[4];
    for (int m = 0; m < 4; m++) h[m].resize(c->esz * n);
    const int comp[4] = {0, 1, 2, 4};
    for (size_t q = 0; q < n; q++) {
        if (cons5[5 * q + 3] != 0.0) return fail(c, EFVM_ERR_UNSUPPORTED, "rho_w must be 0 on the 2-D GPU path");
        for (int m = 0; m < 4; m++) {
            double v = cons5[5 * q + comp[m]];
            if (c->esz == 8) reinterpret_cast<double*>(h[m].data())[q] = v;
            else reinterpret_cast<float*>(h[m].data())[q] = (float)v;
        }
    }
    for (int m = 0; m < 4; m++)
        HIPCHK(c, hipMemcpyAsync(c->P[set][m], h[m].data(), c->esz * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm_download_state(efvm_ctx* c, double* cons5, size_t n) {
    if (!c || !cons5) return EFVM_ERR_ARG;
    if ((long)n != c->ncell) return fail(c, EFVM_ERR_ARG, "download: n_cells != nx*ny");
    // the state of the current substage: A at a step boundary, else B / C
    int set = stage_set(c);
    std::vector<char> h[4];
    for (int m = 0; m < 4; m++) {
        h[m].resize(c->esz * n);
        HIPCHK(c, hipMemcpyAsync(h[m].data(), c->P[set][m], c->esz * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int comp[4] = {0, 1, 2, 4};
    for (size_t q = 0; q < n; q++) {
        cons5[5 * q + 3] = 0.0;
        for (int m = 0; m < 4; m++)
            cons5[5 * q + comp[m]] = c->esz == 8 ? reinterpret_cast<double*>(h[m].data())[q]
                                                 : (double)reinterpret_cast<float*>(h[m].data())[q];
    }
    return EFVM_OK;
}

efvm_status efvm_get_clock(efvm_ctx* c, double* dt, double* time, long* step, int* stage) {
    if (!c) return EFVM_ERR_ARG;
    DevClock h;
    HIPCHK(c, hipMemcpyAsync(&h, c->clock, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (dt) *dt = h.dt;
    if (time) *time = h.time;
    if (step) *step = (long)h.step;
    if (stage) *stage = c->rk_stage;
    if (h.bad) return fail(c, EFVM_ERR_NONPHYSICAL, "CFL scan saw " + std::to_string(h.bad) + " non-finite or non-positive-density cells");
    return EFVM_OK;
}

efvm_status efvm_set_clock(efvm_ctx* c, double time, long step) {
    if (!c) return EFVM_ERR_ARG;
    HIPCHK(c, launch_clock_set(c->clock, 0.0, time, step, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lam_valid = 0; c->rk_stage = 0; c->parity = 0;
    return EFVM_OK;
}

efvm_status efvm_cfl(efvm_ctx* c, double* dt_out) {
    if (!c || !dt_out) return EFVM_ERR_ARG;
    if (c->sch.fixed_dt > 0.0) { *dt_out = c->sch.fixed_dt; return EFVM_OK; }
    efvm_status st = launch_cfl_into(c, c->P[c->cur], &c->clock->lam_tmp);
    if (st != EFVM_OK) return st;
    double lam = 0.0;
    HIPCHK(c, hipMemcpyAsync(&lam, &c->clock->lam_tmp, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *dt_out = lam > 0.0 ? c->sch.cfl / lam : INFINITY;
    return EFVM_OK;
}

efvm_status efvm_run(efvm_ctx* c) {
    if (!c) return EFVM_ERR_ARG;
    return run_substage(c);
}

efvm_status efvm_step(efvm_ctx* c, int n_steps) {
    if (!c || n_steps < 0) return EFVM_ERR_ARG;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    const int nsub = c->sch.integrator == EFVM_INTEGRATOR_RK3 ? 3 : 1;
    for (int k = 0; k < n_steps; k++) {
        // from a step boundary a whole step; from mid-step (after run() substages) three substages
        if (c->rk_stage == 0) {
            efvm_status st = run_step(c);
            if (st != EFVM_OK) return st;
        } else {
            for (int q = 0; q < nsub; q++) {
                efvm_status st = run_substage(c);
                if (st != EFVM_OK) return st;
            }
        }
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timed = 1;
    return EFVM_OK;
}

efvm_status efvm_advance_to(efvm_ctx* c, double t_final, long max_steps, long* done) {
    if (!c) return EFVM_ERR_ARG;
    if (c->rk_stage != 0) return fail(c, EFVM_ERR_ARG, "advance_to must start at a step boundary");
    long k = 0;
    double t = 0.0;
    efvm_status st = efvm_get_clock(c, nullptr, &t, nullptr, nullptr);
    if (st != EFVM_OK && st != EFVM_ERR_NONPHYSICAL) return st;
    // the loop condition needs the clock every step (t >= t_final before each whole step)
    while (t < t_final && k < max_steps) {
        if ((st = run_step(c)) != EFVM_OK) return st;
        k++;
        DevClock h;
        HIPCHK(c, hipMemcpyAsync(&h, c->clock, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        t = h.time;
        if (h.bad) { if (done) *done = k; return fail(c, EFVM_ERR_NONPHYSICAL, "non-physical state during advance"); }
    }
    if (done) *done = k;
    return EFVM_OK;
}

efvm_status efvm_synchronize(efvm_ctx* c) {
    if (!c) return EFVM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm_graph_capture(efvm_ctx* c, int n_steps) {
    if (!c || n_steps <= 0 || (n_steps & 1)) return fail(c, EFVM_ERR_ARG, "graph capture needs an even positive step count");
    if (c->rk_stage != 0) return fail(c, EFVM_ERR_ARG, "graph capture must start at a step boundary");
    drop_graph(c);
    // dt of the first captured step must already be on the device
    if (!c->lam_valid && c->sch.fixed_dt <= 0.0) {
        efvm_status st = launch_cfl_into(c, c->P[c->cur], &c->clock->lam[c->parity]);
        if (st != EFVM_OK) return st;
        c->lam_valid = 1;
    }
    if (c->n_cyl > 0 && !c->ibm_flipped) {
        HIPCHK(c, launch_ibm_flip(c->P[c->cur], (int)c->esz, c->solid, c->ncell, c->stagnant, c->stream));
        c->ibm_flipped = 1;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int cur0 = c->cur, prev0 = c->prev, par0 = c->parity, have0 = c->have_prev;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    efvm_status st = EFVM_OK;
    for (int k = 0; k < n_steps && st == EFVM_OK; k++) st = run_step(c);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st != EFVM_OK) { if (g) (void)hipGraphDestroy(g); return st; }
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, std::string("end capture: ") + hipGetErrorString(e));
    c->graph = g;
    HIPCHK(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
    c->graph_steps = n_steps;
    // capture only recorded work: restore the host-side step machinery (an even step count
    // returns both the CFL slot parity and the Euler ping-pong to their starting values)
    c->cur = cur0; c->prev = prev0; c->parity = par0; c->rk_stage = 0; c->have_prev = have0;
    c->graph_cur = cur0; c->graph_parity = par0;
    return EFVM_OK;
}

efvm_status efvm_graph_launch(efvm_ctx* c) {
    if (!c || !c->graph_exec) return fail(c, EFVM_ERR_ARG, "no captured graph");
    // The captured steps read U^n from the plane set and lambda_max from the CFL slot they were
    // captured at.  Eager work since then may have left the context mid-step (a lone run()
    // substage), on the other Euler plane set or CFL slot (an odd number of eager steps), or
    // with a new state whose CFL maximum is not computed yet (an upload or a checkpoint load).
    if (c->rk_stage != 0) return fail(c, EFVM_ERR_ARG, "graph replay must start at a step boundary");
    if (c->cur != c->graph_cur || c->parity != c->graph_parity)
        return fail(c, EFVM_ERR_ARG, "the state is on another plane set or CFL slot than at capture (an odd number of "
                                     "eager steps since): capture the graph again");
    if (!c->lam_valid && c->sch.fixed_dt <= 0.0) {
        efvm_status st = launch_cfl_into(c, c->P[c->cur], &c->clock->lam[c->parity]);
        if (st != EFVM_OK) return st;
        c->lam_valid = 1;
    }
    if (c->n_cyl > 0 && !c->ibm_flipped) {
        HIPCHK(c, launch_ibm_flip(c->P[c->cur], (int)c->esz, c->solid, c->ncell, c->stagnant, c->stream));
        c->ibm_flipped = 1;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timed = 1;
    c->have_prev = 1;             // the replayed steps end with the roles they started with
    c->ghost_hist = 1;
    return EFVM_OK;
}

efvm_status efvm_last_elapsed_ms(efvm_ctx* c, double* ms) {
    if (!c || !ms || !c->timed) return EFVM_ERR_ARG;
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return EFVM_OK;
}

efvm_status efvm_ibm_add_cylinder(efvm_ctx* c, double cx, double cy, double r, const double fs[5]) {
    if (!c || !fs || !(r > 0.0)) return EFVM_ERR_ARG;
    if (c->n_cyl >= 8) return fail(c, EFVM_ERR_UNSUPPORTED, "at most 8 cylinders");
    c->cyl[c->n_cyl][0] = cx; c->cyl[c->n_cyl][1] = cy; c->cyl[c->n_cyl][2] = r;
    c->n_cyl++;
    std::memcpy(c->freestream, fs, sizeof(double) * 5);
    drop_graph(c);
    free_ibm(c);
    // mask, ghost list and markers on the host in FP64, exactly as the reference
    // builds them (cfd_ibm:155-254, 566-598); a static body needs them once.
    const efvm_grid& g = c->grid;
    const long n = c->ncell;
    double dx = (g.hi_x - g.lo_x) / g.nx, dy = (g.hi_y - g.lo_y) / g.ny, dz = g.dz;
    std::vector<uint8_t> solid(n, 0);
    auto sd = [&](int b, double x, double y) {
        double ddx = x - c->cyl[b][0], ddy = y - c->cyl[b][1];
        return std::sqrt(ddx * ddx + ddy * ddy) - c->cyl[b][2];
    };
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
            double x = g.lo_x + ((double)i + 0.5) * dx, y = g.lo_y + ((double)j + 0.5) * dy;
            for (int b = 0; b < c->n_cyl; b++) if (sd(b, x, y) < 0.0) solid[(long)j * g.nx + i] = 1;
        }
    std::vector<int> gidx;
    std::vector<double> gimg;
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
            long q = (long)j * g.nx + i;
            if (!solid[q]) continue;
            // z neighbours are slip-wall ghosts (never solid) on the one-cell-thick mesh,
            // so every solid cell has a fluid neighbour and is an IBM ghost.
            double x = g.lo_x + ((double)i + 0.5) * dx, y = g.lo_y + ((double)j + 0.5) * dy, z = 0.0 + 0.5 * dz;
            double best_abs = 1.79769313486231570815e308, best_signed = 0.0, bn[3] = {1.0, 0.0, 0.0};
            for (int b = 0; b < c->n_cyl; b++) {
                double s = sd(b, x, y);
                if (std::fabs(s) < best_abs) {
                    best_abs = std::fabs(s); best_signed = s;
                    double ddx = x - c->cyl[b][0], ddy = y - c->cyl[b][1];
                    double rr = std::sqrt(ddx * ddx + ddy * ddy);
                    if (rr < 1e-12) { bn[0] = 1.0; bn[1] = 0.0; } else { bn[0] = ddx / rr; bn[1] = ddy / rr; }
                    bn[2] = 0.0;
                }
            }
            double push = 2.0 * (-best_signed);
            gidx.push_back((int)q);
            gimg.push_back(x + push * bn[0]); gimg.push_back(y + push * bn[1]); gimg.push_back(z + push * bn[2]);
        }
    std::vector<double> mk;
    const double PI = 3.14159265358979323846;
    double z_lo = (0.0 + 0.5 * dz) - 0.5 * dz, z_hi = (0.0 + 0.5 * dz) + 0.5 * dz, lz = z_hi - z_lo;
    for (int b = 0; b < c->n_cyl; b++) {
        double area = 2.0 * PI * c->cyl[b][2] * lz / (double)(64 * 1);
        for (int it = 0; it < 64; it++) {
            double th = 2.0 * PI * ((double)it + 0.5) / 64.0;
            double nxv = std::cos(th), nyv = std::sin(th);
            double z = z_lo + (0.0 + 0.5) * lz / 1.0;
            double m[7] = {c->cyl[b][0] + c->cyl[b][2] * nxv, c->cyl[b][1] + c->cyl[b][2] * nyv, z, nxv, nyv, 0.0, area};
            mk.insert(mk.end(), m, m + 7);
        }
    }
    C5<double> f = {fs[0], fs[1], fs[2], fs[3], fs[4]};
    double stag[5] = {f.r, 0.0, 0.0, 0.0, f.re - ref::kinetic(f)};
    efvm_status st = EFVM_OK;
    if ((st = device_alloc(c, (void**)&c->solid, n)) != EFVM_OK) return st;
    if ((st = device_alloc(c, (void**)&c->ghost_idx, sizeof(int) * std::max<size_t>(1, gidx.size()))) != EFVM_OK) return st;
    if ((st = device_alloc(c, (void**)&c->ghost_img, sizeof(double) * std::max<size_t>(3, gimg.size()))) != EFVM_OK) return st;
    for (int k = 0; k < 2; k++)
        if ((st = device_alloc(c, &c->ghost_pre[k], 4 * c->esz * std::max<size_t>(1, gidx.size()))) != EFVM_OK) return st;
    if ((st = device_alloc(c, (void**)&c->markers, sizeof(double) * mk.size())) != EFVM_OK) return st;
    if ((st = device_alloc(c, (void**)&c->stagnant, sizeof(double) * 5)) != EFVM_OK) return st;
    HIPCHK(c, hipMemcpy(c->solid, solid.data(), n, hipMemcpyHostToDevice));
    if (!gidx.empty()) {
        HIPCHK(c, hipMemcpy(c->ghost_idx, gidx.data(), sizeof(int) * gidx.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->ghost_img, gimg.data(), sizeof(double) * gimg.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemcpy(c->markers, mk.data(), sizeof(double) * mk.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->stagnant, stag, sizeof(stag), hipMemcpyHostToDevice));
    c->n_ghost = (int)gidx.size();
    c->n_mark = (int)(mk.size() / 7);
    c->ibm_flipped = 0;
    c->ghost_hist = 0;
    return EFVM_OK;
}

efvm_status efvm_marker_force(efvm_ctx* c, double out[6]) {
    if (!c || !out) return EFVM_ERR_ARG;
    if (c->n_cyl == 0) { for (int d = 0; d < 6; d++) out[d] = 0.0; return EFVM_OK; }
    StageArgs a = base_args(c);
    int set = stage_set(c);
    set_planes(a, c->P[set], c->P[set], c->P[set]);
    IbmArgs ib = ibm_args(c, a);
    double* dout = c->scratch + 4096;
    double dx = a.dx;
    hipError_t e = c->esz == 8 ? launch_marker_force<double>(a, ib, c->n_mark, c->markers, dx, a.visc_const, a.mu, dout, c->stream)
                               : launch_marker_force<float>(a, ib, c->n_mark, c->markers, dx, a.visc_const, a.mu, dout, c->stream);
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, "marker force");
    HIPCHK(c, hipMemcpyAsync(out, dout, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm_wall_force(efvm_ctx* c, double out[6]) {
    if (!c || !out) return EFVM_ERR_ARG;
    // the mask takes the body on the first substage (the reference's Setup phase): no solid cell before
    if (c->n_cyl == 0 || !c->ibm_flipped) { for (int d = 0; d < 6; d++) out[d] = 0.0; return EFVM_OK; }
    // Ghost cells: fills of the substage inputs since the step boundary, combined by the updates.
    // Mid-step U^n is cur and U1 prev; after a step U^n is prev and U2 the third set.
    const bool rk3 = c->sch.integrator == EFVM_INTEGRATOR_RK3;
    const int free_set = 3 - c->cur - c->prev;
    int n_src = 0, s0 = 0, s1 = 0;
    double w0 = 1.0, w1 = 0.0;
    if (c->rk_stage == 1) { n_src = 1; s0 = c->cur; }
    else if (c->rk_stage == 2) { n_src = 2; s0 = c->cur; s1 = c->prev; w0 = 0.75; w1 = 0.25; }
    else if (c->ghost_hist && !rk3) { n_src = 1; s0 = c->prev; }
    else if (c->ghost_hist) { n_src = 2; s0 = c->prev; s1 = free_set; w0 = 1.0 / 3.0; w1 = 2.0 / 3.0; }
    const int set = stage_set(c);
    StageArgs a = base_args(c);
    set_planes(a, c->P[set], c->P[set], c->P[set]);
    IbmArgs ib = ibm_args(c, a);
    double* dout = c->scratch + 4096;
    void* const* p0 = c->P[s0];
    void* const* p1 = c->P[s1];
    hipError_t e = c->esz == 8
        ? launch_wall_force<double>(a, ib, c->n_ghost, c->ghost_idx, n_src, p0, p1, c->ghost_pre[0], c->ghost_pre[1], w0, w1,
                                    a.visc_const, a.mu, dout, c->stream)
        : launch_wall_force<float>(a, ib, c->n_ghost, c->ghost_idx, n_src, p0, p1, c->ghost_pre[0], c->ghost_pre[1], w0, w1,
                                   a.visc_const, a.mu, dout, c->stream);
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, "wall force");
    HIPCHK(c, hipMemcpyAsync(out, dout, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm_conservation_sums(efvm_ctx* c, double out[5]) {
    if (!c || !out) return EFVM_ERR_ARG;
    StageArgs a = base_args(c);
    set_planes(a, c->P[c->cur], c->P[c->cur], c->P[c->cur]);
    double vol = a.dx * a.dy * a.dz;
    double* dout = c->scratch + 4096 + 8;
    int nb = (int)std::min<long>(1024, (c->ncell + 255) / 256);
    hipError_t e = c->esz == 8 ? launch_sums<double>(a, c->scratch, nb, vol, dout, c->stream)
                               : launch_sums<float>(a, c->scratch, nb, vol, dout, c->stream);
    if (e != hipSuccess) return fail(c, EFVM_ERR_HIP, "sums");
    double h[4];
    HIPCHK(c, hipMemcpyAsync(h, dout, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = h[0]; out[1] = h[1]; out[2] = h[2]; out[3] = 0.0; out[4] = h[3];
    return EFVM_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free conservation sums.  The per-block partials live in a fixed device array
// (each GPU holds its own copy for the life of the process); its address is taken once per
// device with that device current.  Host code uses only that address, never the array's name.
// ---------------------------------------------------------------------------
}  // extern "C"
namespace {
constexpr int NOALLOC_BLOCKS = 1024;                     // partial-sum blocks (4 doubles each)
static __device__ __align__(256) double scratch[NOALLOC_BLOCKS * 4];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

__global__ void sums5_final_kernel(const double* __restrict__ part, int nb, double vol, double* out5) {
    if (threadIdx.x != 0) return;
    double s[4] = {0, 0, 0, 0};
    for (int b = 0; b < nb; b++) for (int m = 0; m < 4; m++) s[m] += part[4 * b + m];
    out5[0] = s[0] * vol; out5[1] = s[1] * vol; out5[2] = s[2] * vol; out5[3] = 0.0; out5[4] = s[3] * vol;
}

// The device address of `scratch` on the current device (taken once per device).
hipError_t scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace
extern "C" {

efvm_status efvm_state_sums_noalloc(const void* const planes[4], int precision, long long n_cells, double cell_volume,
                                    double* out5, void* stream) {
    if (!planes || !out5 || !planes[0] || !planes[1] || !planes[2] || !planes[3]) return EFVM_ERR_ARG;
    if (precision != EFVM_FP32 && precision != EFVM_FP64) return EFVM_ERR_ARG;
    if (n_cells < 1 || n_cells > (long long)EFVM_NOALLOC_MAX_CELLS) return EFVM_ERR_ARG;
    double* part = nullptr;
    if (scratch_ptr(&part) != hipSuccess) return EFVM_ERR_HIP;
    StageArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nx = (int)n_cells;
    a.ny = 1;
    for (int m = 0; m < 4; m++) a.uin[m] = planes[m];
    const hipStream_t st = (hipStream_t)stream;
    const int nb = (int)std::min<long long>(NOALLOC_BLOCKS, (n_cells + 255) / 256);
    // every partial the final pass reads is written by this call: nothing to clear
    hipError_t e = precision == EFVM_FP64 ? launch_sums_partial<double>(a, part, nb, st)
                                          : launch_sums_partial<float>(a, part, nb, st);
    if (e != hipSuccess) return EFVM_ERR_HIP;
    sums5_final_kernel<<<1, 64, 0, st>>>(part, nb, cell_volume, out5);
    return hipGetLastError() == hipSuccess ? EFVM_OK : EFVM_ERR_HIP;
}

efvm_status efvm_device_state(efvm_ctx* c, void* planes[4]) {
    if (!c || !planes) return EFVM_ERR_ARG;
    for (int m = 0; m < 4; m++) planes[m] = c->P[c->cur][m];
    return EFVM_OK;
}

}  // extern "C"

// ===========================================================================
// Batched ensembles: hundreds of independent 2-D cases in one persistent launch
// per RK stage (FP32, MUSCL/first-order, any flux, RK3), per-case dt.
// ===========================================================================
namespace efvm {
hipError_t launch_stage_f32_batch(const BatchCase* cases, const StageArgs& rep, const int2* work, int n_work, int stage,
                                  int parity, int blocks, int waves, hipStream_t st);
hipError_t launch_stage_f32_batch_pair(const BatchCase* cases, const StageArgs& rep, const int2* work, int n_work, int stage,
                                       int parity, int blocks, int waves, hipStream_t st);
hipError_t launch_stage_f32_batch_pair23(const BatchCase* cases, const StageArgs& rep, const int2* work, int n_work,
                                         int parity, hipStream_t st);
}

struct efvm_batch {
    int n = 0;
    efvm_scheme sch{};
    std::vector<efvm_case> cases;
    std::vector<BatchCase> hc;
    std::vector<int> strip_rows;
    BatchCase* dc = nullptr;
    int2* dwork = nullptr;         // (case, strip) work items of the one-column kernel
    int n_work = 0;
    int2* dwork_pair = nullptr;    // and of the two-column kernel (stage 1 of every case it takes)
    int n_work_pair = 0;
    int2* dwork_pair_nf = nullptr; // the two-column cases that are not fused (their stages 2 and 3)
    int n_work_pair_nf = 0;
    int2* dwork_fused = nullptr;   // (case, strip) items of the fused stage-2+3 launch
    int n_work_fused = 0;
    int allow_pair = 1;            // cases the two-column kernel takes run on it
    int allow_fuse = 1;            // ... and the periodic ones run stages 2 and 3 fused
    int strip_h23 = 37;            // rows per strip of the fused launch
    DevClock* clocks = nullptr;
    double* lam_slots = nullptr;   // per-case regions of per-strip CFL maxima (the last stage's
    long n_lam_slots = 0;          // strips write them, batch_lam_fold_kernel folds them)
    float* pool = nullptr;
    std::vector<long> offset;      // per-case start (cells) in the pool
    long total_cells = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int timed = 0;
    int parity = 0;
    int lam_valid = 0;
    int blocks = 0, waves = 4;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_parity = 0;          // CFL slot the captured steps start from
    void* snap = nullptr;
    int snap_parity = 0, snap_lam = 0;
    std::string err;
};

namespace {
efvm_status bfail(efvm_batch* b, efvm_status s, const std::string& m) {
    if (b) b->err = m;
    return s;
}
#define BHIPCHK(b, expr)                                                                                 \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return bfail(b, EFVM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

void batch_drop_graph(efvm_batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr; b->graph = nullptr;
}

// Stage arguments of case i (the same builder as a single context, no viscous/IBM).
StageArgs batch_case_args(const efvm_batch* b, int i) {
    efvm_ctx tmp;
    const efvm_case& c = b->cases[i];
    tmp.grid = {c.nx, c.ny, c.lo_x, c.lo_y, c.hi_x, c.hi_y, c.dz};
    tmp.gas = {c.gamma, 287.058, EFVM_VISC_SUTHERLAND, 0.0};
    tmp.sch = b->sch;
    for (int s = 0; s < 4; s++) tmp.bc[s] = c.sides[s];
    tmp.clock = b->clocks + i;
    tmp.strip_h = b->strip_rows[i];
    tmp.xcd_remap = 0;
    tmp.esz = 4;
    // the pool holds plane m of set s of every case at (4 s + m) * total_cells: one plane stride
    tmp.plane_stride = (size_t)b->total_cells * 4 <= 0xFFFFFFFFu ? (size_t)b->total_cells * 4 : 0;
    tmp.allow_pair = b->allow_pair;
    StageArgs a = base_args(&tmp);
    a.fuse_cfl = b->sch.fixed_dt <= 0.0 ? 1 : 0;
    return a;
}

// The work table, XCD-aware.  Workgroup k of a launch runs on XCD k % 8 and each XCD has its own
// L2, so a strip shares the rows it reads beyond its own (two above, two below) with its
// vertical neighbours only when they run on the same XCD.  Each case's strips, in row-major
// order (a band of whole strip rows at a time), are dealt to per-XCD queues: a case of fewer
// strips than one queue's share goes whole to the least-loaded queue, a larger one is cut into
// contiguous bands, one per queue.  Cases go largest first, so each queue ends on small work.
// Position p of the table belongs to workgroup p / waves, hence to XCD (p / waves) % 8, and is
// filled from that XCD's queue (or, once it is empty, from the longest remaining one).
// kind: 0 the one-column cases, 1 every two-column case (stage 1), 2 the two-column cases that are
// not fused (stages 2 and 3), 3 the fused cases in the fused launch's strips (one workgroup per item).
std::vector<int2> batch_work_table(const efvm_batch* b, int kind) {
    constexpr int NXCD = 8;
    std::vector<int> order;
    for (int i = 0; i < b->n; i++) {
        const bool pair = b->hc[i].a.pair_path != 0, fused = b->hc[i].fused != 0;
        if ((kind == 0 && !pair) || (kind == 1 && pair) || (kind == 2 && pair && !fused) || (kind == 3 && fused))
            order.push_back(i);
    }
    auto cells = [&](int i) { return (long)b->cases[i].nx * b->cases[i].ny; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cells(x) > cells(y); });
    std::vector<std::vector<int2>> q(NXCD);
    std::vector<long> load(NXCD, 0);
    long total = 0;
    for (int i : order) total += cells(i);
    const long share = std::max<long>(1, total / NXCD);
    for (int i : order) {
        const StageArgs& a = kind == 3 ? b->hc[i].a23 : b->hc[i].a;
        const int ns = a.n_strips_x * a.n_strips_y;
        const long per = std::max<long>(1, cells(i) / ns);
        if (cells(i) <= share / 4 || ns < NXCD) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            for (int s = 0; s < ns; s++) q[k].push_back(make_int2(i, s));
            load[k] += cells(i);
        } else {
            for (int k = 0; k < NXCD; k++) {
                const int s0 = (int)((long)ns * k / NXCD), s1 = (int)((long)ns * (k + 1) / NXCD);
                for (int s = s0; s < s1; s++) q[k].push_back(make_int2(i, s));
                load[k] += per * (s1 - s0);
            }
        }
    }
    size_t n_items = 0;
    for (auto& v : q) n_items += v.size();
    std::vector<size_t> head(NXCD, 0);
    std::vector<int2> work;
    work.reserve(n_items);
    const int W = kind == 3 ? 1 : std::max(1, b->waves);
    while (work.size() < n_items) {
        const int k = (int)((work.size() / W) % NXCD);
        int src = k;
        if (head[src] == q[src].size()) {
            size_t best = 0;
            for (int t = 0; t < NXCD; t++)
                if (q[t].size() - head[t] > best) { best = q[t].size() - head[t]; src = t; }
        }
        work.push_back(q[src][head[src]++]);
    }
    return work;
}

// The set holding case i's U^n: A, or for a fused case at odd step parity C.
int batch_cur_set(const efvm_batch* b, int i) { return b->hc[i].fused && (b->parity & 1) ? 2 : 0; }

efvm_status batch_rebuild(efvm_batch* b) {
    batch_drop_graph(b);
    // a case's U^n moves between sets A and C when it becomes fused or stops being fused at odd
    // parity: bring it back to A first (stream order keeps it behind the steps enqueued before)
    std::vector<int> was_cur(b->n);
    for (int i = 0; i < b->n; i++) was_cur[i] = batch_cur_set(b, i);
    for (int i = 0; i < b->n; i++) {
        BatchCase& bc = b->hc[i];
        bc.a = batch_case_args(b, i);
        for (int s = 0; s < 3; s++)
            for (int m = 0; m < 4; m++) bc.P[s][m] = b->pool + (long)(4 * s + m) * b->total_cells + b->offset[i];
        bc.fused = b->allow_fuse && bc.a.pair_path == 1 && bc.a.ny >= 8 ? 1 : 0;
        bc.a23 = bc.a;
        bc.a23.strip_h = b->strip_h23;
        bc.a23.n_strips_x = (bc.a.nx + STRIP_OUT23 - 1) / STRIP_OUT23;
        bc.a23.n_strips_y = (bc.a.ny + b->strip_h23 - 1) / b->strip_h23;
    }
    // CFL slots: one per strip of the launch that runs the case's last stage (same-address
    // atomics from thousands of strips of one large case slowed that stage by a quarter)
    {
        long need = 0;
        std::vector<long> off(b->n);
        for (int i = 0; i < b->n; i++) {
            const BatchCase& bc = b->hc[i];
            off[i] = need;
            need += std::max<long>((long)bc.a.n_strips_x * bc.a.n_strips_y, (long)bc.a23.n_strips_x * bc.a23.n_strips_y);
        }
        if (need > b->n_lam_slots) {
            if (b->lam_slots) (void)hipFree(b->lam_slots);
            b->lam_slots = nullptr;
            b->n_lam_slots = 0;
            BHIPCHK(b, hipMalloc((void**)&b->lam_slots, sizeof(double) * need));
            b->n_lam_slots = need;
        }
        for (int i = 0; i < b->n; i++) {
            b->hc[i].a.lam_part = b->lam_slots + off[i];
            b->hc[i].a23.lam_part = b->lam_slots + off[i];
        }
    }
    for (int i = 0; i < b->n; i++) {
        const int now = batch_cur_set(b, i);
        if (now != was_cur[i]) {
            const size_t bytes = sizeof(float) * b->cases[i].nx * b->cases[i].ny;
            for (int m = 0; m < 4; m++)
                BHIPCHK(b, hipMemcpyAsync(b->hc[i].P[now][m], b->hc[i].P[was_cur[i]][m], bytes, hipMemcpyDeviceToDevice, b->stream));
        }
    }
    const std::vector<int2> work = batch_work_table(b, 0), work_pair = batch_work_table(b, 1),
                            work_pair_nf = batch_work_table(b, 2), work_fused = batch_work_table(b, 3);
    struct T { int2** d; int* n; const std::vector<int2>* v; };
    for (const T& t : {T{&b->dwork, &b->n_work, &work}, T{&b->dwork_pair, &b->n_work_pair, &work_pair},
                       T{&b->dwork_pair_nf, &b->n_work_pair_nf, &work_pair_nf},
                       T{&b->dwork_fused, &b->n_work_fused, &work_fused}}) {
        if (*t.d) (void)hipFree(*t.d);
        *t.d = nullptr;
        BHIPCHK(b, hipMalloc((void**)t.d, sizeof(int2) * std::max<size_t>(1, t.v->size())));
        if (!t.v->empty())
            BHIPCHK(b, hipMemcpyAsync(*t.d, t.v->data(), sizeof(int2) * t.v->size(), hipMemcpyHostToDevice, b->stream));
        *t.n = (int)t.v->size();
    }
    BHIPCHK(b, hipMemcpyAsync(b->dc, b->hc.data(), sizeof(BatchCase) * b->n, hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return EFVM_OK;
}

efvm_status batch_cfl_all(efvm_batch* b, int slot) {
    for (int i = 0; i < b->n; i++) {
        StageArgs a = b->hc[i].a;
        hipError_t e = launch_cfl_f32(a, b->hc[i].P[batch_cur_set(b, i)], &b->clocks[i].lam[slot], b->stream);
        if (e != hipSuccess) return bfail(b, EFVM_ERR_HIP, "batch cfl");
    }
    return EFVM_OK;
}

efvm_status batch_enqueue_steps(efvm_batch* b, int n_steps) {
    // kernel features = union over the cases (an EOS-based boundary anywhere selects the general kernel)
    StageArgs rep = b->hc[0].a;
    for (int i = 1; i < b->n; i++)
        for (int s = 0; s < 4; s++) {
            int k = b->hc[i].a.bc_kind[s];
            if (k == BC_SUP_INFLOW || k == BC_SUB_INFLOW || k == BC_PRESSURE_OUTLET) rep.bc_kind[s] = k;
        }
    // 0 = one wave per (case, strip) work item: the hardware dispatcher then balances the
    // uneven strips; a persistent grid striding through the table measured up to 13% slower
    auto nblocks = [&](int n_items) { return b->blocks > 0 ? b->blocks : std::max(1, (n_items + b->waves - 1) / b->waves); };
    for (int k = 0; k < n_steps; k++) {
        for (int st = 0; st < 3; st++) {
            // the kernels of a stage take disjoint cases, so their order does not matter; the fused
            // cases take part in stage 1 only, then run stages 2 and 3 in their own launch
            const int2* wp = st == 0 ? b->dwork_pair : b->dwork_pair_nf;
            const int np = st == 0 ? b->n_work_pair : b->n_work_pair_nf;
            if (np > 0) {
                hipError_t e = launch_stage_f32_batch_pair(b->dc, rep, wp, np, st, b->parity, nblocks(np), b->waves, b->stream);
                if (e != hipSuccess) return bfail(b, EFVM_ERR_HIP, std::string("batch stage: ") + hipGetErrorString(e));
            }
            if (b->n_work > 0) {
                hipError_t e = launch_stage_f32_batch(b->dc, rep, b->dwork, b->n_work, st, b->parity, nblocks(b->n_work),
                                                      b->waves, b->stream);
                if (e != hipSuccess) return bfail(b, EFVM_ERR_HIP, std::string("batch stage: ") + hipGetErrorString(e));
            }
        }
        if (b->n_work_fused > 0) {
            hipError_t e = launch_stage_f32_batch_pair23(b->dc, rep, b->dwork_fused, b->n_work_fused, b->parity, b->stream);
            if (e != hipSuccess) return bfail(b, EFVM_ERR_HIP, std::string("batch fused stages: ") + hipGetErrorString(e));
        }
        if (rep.fuse_cfl) {
            hipError_t e = launch_batch_lam_fold(b->dc, b->n, b->parity, b->stream);
            if (e != hipSuccess) return bfail(b, EFVM_ERR_HIP, std::string("batch cfl fold: ") + hipGetErrorString(e));
        }
        b->parity ^= 1;
    }
    return EFVM_OK;
}
}  // namespace

extern "C" {

efvm_status efvm_batch_create(int n_cases, const efvm_case* cases, const efvm_scheme* scheme, efvm_batch** out) {
    if (!cases || !scheme || !out || n_cases <= 0) return EFVM_ERR_ARG;
    *out = nullptr;
    if (scheme->precision != EFVM_FP32 || scheme->integrator != EFVM_INTEGRATOR_RK3 || scheme->viscous)
        return EFVM_ERR_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EFVM_ERR_NO_DEVICE;
    for (int i = 0; i < n_cases; i++) {
        const efvm_case& c = cases[i];
        if (c.nx < 1 || c.ny < 1 || !(c.hi_x > c.lo_x) || !(c.hi_y > c.lo_y) || !(c.dz > 0.0)) return EFVM_ERR_CONFIG;
        for (int ax = 0; ax < 2; ax++)
            if ((c.sides[2 * ax].kind == EFVM_BC_PERIODIC) != (c.sides[2 * ax + 1].kind == EFVM_BC_PERIODIC)) return EFVM_ERR_CONFIG;
        for (int s = 0; s < 4; s++)
            if (c.sides[s].kind < 0 || c.sides[s].kind > EFVM_BC_PRESSURE_OUTLET) return EFVM_ERR_UNSUPPORTED;   // no ghost tables
    }
    efvm_batch* b = new (std::nothrow) efvm_batch();
    if (!b) return EFVM_ERR_ARG;
    b->n = n_cases;
    b->sch = *scheme;
    b->cases.assign(cases, cases + n_cases);
    b->hc.resize(n_cases);
    b->strip_rows.assign(n_cases, 31);
    b->offset.resize(n_cases);
    long tot = 0;
    for (int i = 0; i < n_cases; i++) {
        b->offset[i] = tot;
        tot += (long)cases[i].nx * cases[i].ny;
        tot = (tot + 63) & ~63L;   // 256-byte aligned case planes
    }
    b->total_cells = tot;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
        hipEventCreate(&b->ev1) != hipSuccess || hipMalloc((void**)&b->pool, sizeof(float) * 12 * tot) != hipSuccess ||
        hipMalloc((void**)&b->clocks, sizeof(DevClock) * n_cases) != hipSuccess ||
        hipMalloc((void**)&b->dc, sizeof(BatchCase) * n_cases) != hipSuccess) {
        efvm_batch_destroy(b);
        return EFVM_ERR_HIP;
    }
    (void)hipMemsetAsync(b->pool, 0, sizeof(float) * 12 * tot, b->stream);
    for (int i = 0; i < n_cases; i++) (void)launch_clock_set(b->clocks + i, 0.0, 0.0, 0, b->stream);
    efvm_status st = batch_rebuild(b);
    if (st != EFVM_OK) { efvm_batch_destroy(b); return st; }
    *out = b;
    return EFVM_OK;
}

void efvm_batch_destroy(efvm_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    if (b->pool) (void)hipFree(b->pool);
    if (b->clocks) (void)hipFree(b->clocks);
    if (b->lam_slots) (void)hipFree(b->lam_slots);
    if (b->dc) (void)hipFree(b->dc);
    if (b->dwork) (void)hipFree(b->dwork);
    if (b->dwork_pair) (void)hipFree(b->dwork_pair);
    if (b->dwork_pair_nf) (void)hipFree(b->dwork_pair_nf);
    if (b->dwork_fused) (void)hipFree(b->dwork_fused);
    if (b->snap) (void)hipFree(b->snap);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* efvm_batch_last_error(const efvm_batch* b) { return b ? b->err.c_str() : "null batch"; }

efvm_status efvm_batch_set_launch(efvm_batch* b, int case_id, int strip_rows) {
    if (!b || case_id >= b->n || strip_rows < 1 || strip_rows > 4096) return bfail(b, EFVM_ERR_ARG, "bad launch shape");
    if (case_id < 0) std::fill(b->strip_rows.begin(), b->strip_rows.end(), strip_rows);
    else b->strip_rows[case_id] = strip_rows;
    return batch_rebuild(b);
}

efvm_status efvm_batch_set_rows(efvm_batch* b, const int* rows) {
    if (!b || !rows) return EFVM_ERR_ARG;
    for (int i = 0; i < b->n; i++)
        if (rows[i] < 1 || rows[i] > 4096) return bfail(b, EFVM_ERR_ARG, "bad launch shape");
    b->strip_rows.assign(rows, rows + b->n);
    return batch_rebuild(b);
}

efvm_status efvm_batch_set_pair_kernel(efvm_batch* b, int allow) {
    if (!b || allow < 0 || allow > 1) return EFVM_ERR_ARG;
    b->allow_pair = allow;
    return batch_rebuild(b);
}

efvm_status efvm_batch_set_fused(efvm_batch* b, int allow, int strip_rows) {
    if (!b || allow < 0 || allow > 1 || strip_rows < 0 || strip_rows > 4096) return bfail(b, EFVM_ERR_ARG, "bad fused setting");
    b->allow_fuse = allow;
    if (strip_rows) b->strip_h23 = strip_rows;
    return batch_rebuild(b);
}

efvm_status efvm_batch_get_fused(const efvm_batch* b, int* allow, int* strip_rows, int* n_fused) {
    if (!b) return EFVM_ERR_ARG;
    if (allow) *allow = b->allow_fuse;
    if (strip_rows) *strip_rows = b->strip_h23;
    if (n_fused) {
        int k = 0;
        for (int i = 0; i < b->n; i++) k += b->hc[i].fused;
        *n_fused = k;
    }
    return EFVM_OK;
}

efvm_status efvm_batch_set_grid(efvm_batch* b, int blocks, int waves_per_block) {
    if (!b || blocks < 0 || waves_per_block < 1 || waves_per_block > 4) return bfail(b, EFVM_ERR_ARG, "bad grid");
    b->blocks = blocks; b->waves = waves_per_block;
    return batch_rebuild(b);   // the work table's XCD order depends on the waves per workgroup
}

efvm_status efvm_batch_upload(efvm_batch* b, int i, const double* cons5, size_t n) {
    if (!b || !cons5 || i < 0 || i >= b->n) return EFVM_ERR_ARG;
    long nc = (long)b->cases[i].nx * b->cases[i].ny;
    if ((long)n != nc) return bfail(b, EFVM_ERR_ARG, "batch upload: n_cells mismatch");
    std::vector<float> h(4 * nc);
    const int comp[4] = {0, 1, 2, 4};
    for (long q = 0; q < nc; q++) {
        if (cons5[5 * q + 3] != 0.0) return bfail(b, EFVM_ERR_UNSUPPORTED, "rho_w must be 0");
        for (int m = 0; m < 4; m++) h[m * nc + q] = (float)cons5[5 * q + comp[m]];
    }
    const int cur = batch_cur_set(b, i);
    for (int m = 0; m < 4; m++)
        BHIPCHK(b, hipMemcpyAsync(b->hc[i].P[cur][m], h.data() + m * nc, sizeof(float) * nc, hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    b->lam_valid = 0;
    return EFVM_OK;
}

efvm_status efvm_batch_download(efvm_batch* b, int i, double* cons5, size_t n) {
    if (!b || !cons5 || i < 0 || i >= b->n) return EFVM_ERR_ARG;
    long nc = (long)b->cases[i].nx * b->cases[i].ny;
    if ((long)n != nc) return bfail(b, EFVM_ERR_ARG, "batch download: n_cells mismatch");
    std::vector<float> h(4 * nc);
    const int cur = batch_cur_set(b, i);
    for (int m = 0; m < 4; m++)
        BHIPCHK(b, hipMemcpyAsync(h.data() + m * nc, b->hc[i].P[cur][m], sizeof(float) * nc, hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    const int comp[4] = {0, 1, 2, 4};
    for (long q = 0; q < nc; q++) {
        cons5[5 * q + 3] = 0.0;
        for (int m = 0; m < 4; m++) cons5[5 * q + comp[m]] = h[m * nc + q];
    }
    return EFVM_OK;
}

efvm_status efvm_batch_step(efvm_batch* b, int n_steps) {
    if (!b || n_steps < 0) return EFVM_ERR_ARG;
    if (!b->lam_valid && b->sch.fixed_dt <= 0.0) {
        efvm_status st = batch_cfl_all(b, b->parity);
        if (st != EFVM_OK) return st;
        b->lam_valid = 1;
    }
    BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
    efvm_status st = batch_enqueue_steps(b, n_steps);
    if (st != EFVM_OK) return st;
    BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
    b->timed = 1;
    return EFVM_OK;
}

efvm_status efvm_batch_graph_capture(efvm_batch* b, int n_steps) {
    if (!b || n_steps <= 0 || (n_steps & 1)) return bfail(b, EFVM_ERR_ARG, "graph capture needs an even positive step count");
    batch_drop_graph(b);
    if (!b->lam_valid && b->sch.fixed_dt <= 0.0) {
        efvm_status st = batch_cfl_all(b, b->parity);
        if (st != EFVM_OK) return st;
        b->lam_valid = 1;
    }
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    const int par0 = b->parity;
    BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    efvm_status st = batch_enqueue_steps(b, n_steps);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    b->parity = par0;
    if (st != EFVM_OK) { if (g) (void)hipGraphDestroy(g); return st; }
    if (e != hipSuccess) return bfail(b, EFVM_ERR_HIP, std::string("end capture: ") + hipGetErrorString(e));
    b->graph = g;
    BHIPCHK(b, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    b->graph_parity = par0;
    return EFVM_OK;
}

efvm_status efvm_batch_graph_launch(efvm_batch* b) {
    if (!b || !b->graph_exec) return bfail(b, EFVM_ERR_ARG, "no captured graph");
    // the captured steps read each case's lambda_max from the CFL slot of the capture; an odd
    // number of eager steps since moved it, and an upload needs a fresh CFL scan first
    if (b->parity != b->graph_parity)
        return bfail(b, EFVM_ERR_ARG, "the cases are on another CFL slot than at capture (an odd number of eager "
                                      "steps since): capture the graph again");
    if (!b->lam_valid && b->sch.fixed_dt <= 0.0) {
        efvm_status st = batch_cfl_all(b, b->parity);
        if (st != EFVM_OK) return st;
        b->lam_valid = 1;
    }
    BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
    BHIPCHK(b, hipGraphLaunch(b->graph_exec, b->stream));
    BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
    b->timed = 1;
    return EFVM_OK;
}

efvm_status efvm_batch_clock(efvm_batch* b, int i, double* dt, double* time, long* step) {
    if (!b || i < 0 || i >= b->n) return EFVM_ERR_ARG;
    DevClock h;
    BHIPCHK(b, hipMemcpyAsync(&h, b->clocks + i, sizeof(h), hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    if (dt) *dt = h.dt;
    if (time) *time = h.time;
    if (step) *step = (long)h.step;
    if (h.bad) return bfail(b, EFVM_ERR_NONPHYSICAL, "case " + std::to_string(i) + ": non-physical cells");
    return EFVM_OK;
}

efvm_status efvm_batch_synchronize(efvm_batch* b) {
    if (!b) return EFVM_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return EFVM_OK;
}

efvm_status efvm_batch_last_elapsed_ms(efvm_batch* b, double* ms) {
    if (!b || !ms || !b->timed) return EFVM_ERR_ARG;
    BHIPCHK(b, hipEventSynchronize(b->ev1));
    float f = 0.f;
    BHIPCHK(b, hipEventElapsedTime(&f, b->ev0, b->ev1));
    *ms = f;
    return EFVM_OK;
}

// Device-side snapshot of the state and clocks (op 0 = save, 1 = restore), so the
// autotuner can time real steps without disturbing a running simulation.
efvm_status efvm_batch_snapshot(efvm_batch* b, int op) {
    if (!b || op < 0 || op > 1) return EFVM_ERR_ARG;
    static_assert(sizeof(DevClock) % 8 == 0, "clock layout");
    // sets A and C: U^n of every case is in one of them (C for fused cases at odd parity)
    const size_t sbytes = sizeof(float) * 4 * b->total_cells, cbytes = sizeof(DevClock) * b->n;
    float* setA = b->pool;
    float* setC = b->pool + 8 * b->total_cells;
    if (!b->snap) {
        if (op == 1) return bfail(b, EFVM_ERR_ARG, "no snapshot saved");
        BHIPCHK(b, hipMalloc(&b->snap, 2 * sbytes + cbytes));
    }
    char* s = (char*)b->snap;
    if (op == 0) {
        BHIPCHK(b, hipMemcpyAsync(s, setA, sbytes, hipMemcpyDeviceToDevice, b->stream));
        BHIPCHK(b, hipMemcpyAsync(s + sbytes, setC, sbytes, hipMemcpyDeviceToDevice, b->stream));
        BHIPCHK(b, hipMemcpyAsync(s + 2 * sbytes, b->clocks, cbytes, hipMemcpyDeviceToDevice, b->stream));
        b->snap_parity = b->parity; b->snap_lam = b->lam_valid;
    } else {
        if (b->snap_parity != b->parity) {
            // fused cases move between A and C with the parity: restore under the saved one
            b->parity = b->snap_parity;
        }
        BHIPCHK(b, hipMemcpyAsync(setA, s, sbytes, hipMemcpyDeviceToDevice, b->stream));
        BHIPCHK(b, hipMemcpyAsync(setC, s + sbytes, sbytes, hipMemcpyDeviceToDevice, b->stream));
        BHIPCHK(b, hipMemcpyAsync(b->clocks, s + 2 * sbytes, cbytes, hipMemcpyDeviceToDevice, b->stream));
        b->lam_valid = b->snap_lam;
    }
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return EFVM_OK;
}

efvm_status efvm_snapshot(efvm_ctx* c, int op) {
    if (!c || op < 0 || op > 1) return EFVM_ERR_ARG;
    if (c->rk_stage != 0) return fail(c, EFVM_ERR_ARG, "snapshot at a step boundary only");
    const size_t pbytes = c->esz * c->ncell;
    if (!c->snap) {
        if (op == 1) return fail(c, EFVM_ERR_ARG, "no snapshot saved");
        HIPCHK(c, hipMalloc(&c->snap, 8 * pbytes + sizeof(DevClock)));
    }
    char* s = (char*)c->snap;
    if (op == 1) { c->cur = c->snap_cur; c->prev = c->snap_prev; }
    // U^n and U^{n-1} (the latter is the checkpoint's u0 column)
    for (int m = 0; m < 4; m++) {
        if (op == 0) {
            HIPCHK(c, hipMemcpyAsync(s + m * pbytes, c->P[c->cur][m], pbytes, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(s + (4 + m) * pbytes, c->P[c->prev][m], pbytes, hipMemcpyDeviceToDevice, c->stream));
        } else {
            HIPCHK(c, hipMemcpyAsync(c->P[c->cur][m], s + m * pbytes, pbytes, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->P[c->prev][m], s + (4 + m) * pbytes, pbytes, hipMemcpyDeviceToDevice, c->stream));
        }
    }
    if (op == 0) {
        HIPCHK(c, hipMemcpyAsync(s + 8 * pbytes, c->clock, sizeof(DevClock), hipMemcpyDeviceToDevice, c->stream));
        c->snap_parity = c->parity; c->snap_lam = c->lam_valid; c->snap_flip = c->ibm_flipped;
        c->snap_cur = c->cur; c->snap_prev = c->prev; c->snap_have_prev = c->have_prev;
    } else {
        HIPCHK(c, hipMemcpyAsync(c->clock, s + 8 * pbytes, sizeof(DevClock), hipMemcpyDeviceToDevice, c->stream));
        c->parity = c->snap_parity; c->lam_valid = c->snap_lam; c->ibm_flipped = c->snap_flip;
        c->have_prev = c->snap_have_prev;
        c->ghost_hist = 0;            // the third set and the pre-mirror ghosts are not in the snapshot
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned face reconstruction: network object, training, attachment to a context
// ---------------------------------------------------------------------------
struct efvm_mlp {
    hipStream_t stream = nullptr;
    float* p = nullptr;            // FP32 master parameters
    float* g = nullptr;            // gradient
    float* m = nullptr;            // Adam moments
    float* v = nullptr;
    long adam_t = 0;
    float* X = nullptr;            // samples x 20
    float* Y = nullptr;            // samples x 4
    long n_data = 0, cap = 0;
    float* partial = nullptr;      // per-block gradient partials
    double* loss_partial = nullptr;
    long part_blocks = 0;
    double* losses = nullptr;      // per-step minibatch losses
    long loss_cap = 0;
    mlp::Packed* packed = nullptr;
    int packed_valid = 0;
    int own4 = 0;                  // the packed correction is the 4-tap own-variable form
    float taps[4][4] = {};         // its open-gate taps (mlp::light_taps)
    std::string err;
};

namespace {

efvm_status mfail(efvm_mlp* m, efvm_status s, const std::string& msg) {
    if (m) m->err = msg;
    return s;
}
#define MHIPCHK(mm, expr)                                                                              \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) return mfail(mm, EFVM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// round to nearest bfloat16, ties to even
uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

void put8(uint32_t dst[4], const float vals[8]) {
    for (int d = 0; d < 4; d++) dst[d] = (uint32_t)to_bf16(vals[2 * d]) | ((uint32_t)to_bf16(vals[2 * d + 1]) << 16);
}

// hidden unit that K index (step s, lane group g, element e) stands for between layers
int perm_h(int s, int g, int e) { return 16 * (2 * s + e / 4) + 4 * g + e % 4; }

void pack_weights(const float* p, mlp::Packed& P) {
    using namespace mlp;
    const float* W1 = p + OFF_W1;
    const float* W2 = p + OFF_W2;
    const float* W3 = p + OFF_W3;
    float v8[8];
    for (int t = 0; t < 4; t++)
        for (int ln = 0; ln < 64; ln++) {
            const int o = 16 * t + ln % 16, g = ln / 16;
            for (int e = 0; e < 8; e++) {
                const int k = 8 * g + e;
                v8[e] = k < NIN ? W1[o * NIN + k] : (k == NIN ? p[OFF_B1 + o] : 0.f);   // feature 20 == 1
            }
            put8(P.w1[t][ln], v8);
            for (int s = 0; s < 2; s++) {
                for (int e = 0; e < 8; e++) v8[e] = W2[o * NH + perm_h(s, g, e)];
                put8(P.w2[t][s][ln], v8);
            }
        }
    for (int s = 0; s < 2; s++)
        for (int ln = 0; ln < 64; ln++) {
            const int r = (ln % 16) % 4, g = ln / 16;
            for (int e = 0; e < 8; e++) v8[e] = W3[r * NH + perm_h(s, g, e)];
            put8(P.w3[s][ln], v8);
        }
    std::memcpy(P.b1, p + OFF_B1, sizeof(float) * NH);
    std::memcpy(P.b2, p + OFF_B2, sizeof(float) * NH);
    std::memcpy(P.b3, p + OFF_B3, sizeof(float) * NOUT);
    std::memcpy(P.L, p + OFF_L, sizeof(float) * NOUT * NIN);
    std::memcpy(P.bL, p + OFF_BL, sizeof(float) * NOUT);
    // own-variable 4-tap correction without bias (what the cubic-exact fit produces)?
    P.own4 = 1;
    for (int v = 0; v < NOUT; v++) {
        if (P.bL[v] != 0.f) P.own4 = 0;
        for (int f = 0; f < NIN; f++) {
            const bool tap = f % 4 == v && f / 4 != 2;
            if (!tap && P.L[v][f] != 0.f) P.own4 = 0;
        }
    }
}

efvm_status mlp_repack(efvm_mlp* m) {
    if (m->packed_valid) return EFVM_OK;
    std::vector<float> h(mlp::NPARAM);
    MHIPCHK(m, hipMemcpyAsync(h.data(), m->p, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    mlp::Packed P;
    pack_weights(h.data(), P);
    m->own4 = P.own4;
    mlp::light_taps(P.L, m->taps);
    MHIPCHK(m, hipMemcpyAsync(m->packed, &P, sizeof(P), hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    m->packed_valid = 1;
    return EFVM_OK;
}

efvm_status mlp_reserve_partials(efvm_mlp* m, long blocks) {
    if (blocks <= m->part_blocks) return EFVM_OK;
    if (m->partial) (void)hipFree(m->partial);
    if (m->loss_partial) (void)hipFree(m->loss_partial);
    m->partial = nullptr; m->loss_partial = nullptr; m->part_blocks = 0;
    MHIPCHK(m, hipMalloc((void**)&m->partial, sizeof(float) * mlp::NPARAM * blocks));
    MHIPCHK(m, hipMalloc((void**)&m->loss_partial, sizeof(double) * blocks));
    m->part_blocks = blocks;
    return EFVM_OK;
}

efvm_status mlp_grow(efvm_mlp* m, long need) {
    if (need <= m->cap) return EFVM_OK;
    long cap = std::max(need, 2 * m->cap);
    float *X = nullptr, *Y = nullptr;
    MHIPCHK(m, hipMalloc((void**)&X, sizeof(float) * mlp::NIN * cap));
    MHIPCHK(m, hipMalloc((void**)&Y, sizeof(float) * mlp::NOUT * cap));
    if (m->n_data) {
        MHIPCHK(m, hipMemcpyAsync(X, m->X, sizeof(float) * mlp::NIN * m->n_data, hipMemcpyDeviceToDevice, m->stream));
        MHIPCHK(m, hipMemcpyAsync(Y, m->Y, sizeof(float) * mlp::NOUT * m->n_data, hipMemcpyDeviceToDevice, m->stream));
        MHIPCHK(m, hipStreamSynchronize(m->stream));
    }
    if (m->X) (void)hipFree(m->X);
    if (m->Y) (void)hipFree(m->Y);
    m->X = X; m->Y = Y; m->cap = cap;
    return EFVM_OK;
}

}  // namespace

extern "C" {

efvm_status efvm_mlp_create(efvm_mlp** out) {
    if (!out) return EFVM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EFVM_ERR_NO_DEVICE;
    efvm_mlp* m = new (std::nothrow) efvm_mlp();
    if (!m) return EFVM_ERR_ARG;
    const size_t pb = sizeof(float) * mlp::NPARAM;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess || hipMalloc((void**)&m->p, pb) != hipSuccess ||
        hipMalloc((void**)&m->g, pb) != hipSuccess || hipMalloc((void**)&m->m, pb) != hipSuccess ||
        hipMalloc((void**)&m->v, pb) != hipSuccess || hipMalloc((void**)&m->packed, sizeof(mlp::Packed)) != hipSuccess) {
        efvm_mlp_destroy(m);
        return EFVM_ERR_HIP;
    }
    if (hipMemsetAsync(m->p, 0, pb, m->stream) != hipSuccess || hipMemsetAsync(m->m, 0, pb, m->stream) != hipSuccess ||
        hipMemsetAsync(m->v, 0, pb, m->stream) != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess) {
        efvm_mlp_destroy(m);
        return EFVM_ERR_HIP;
    }
    *out = m;
    return EFVM_OK;
}

void efvm_mlp_destroy(efvm_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* q : {(void*)m->p, (void*)m->g, (void*)m->m, (void*)m->v, (void*)m->X, (void*)m->Y, (void*)m->partial,
                    (void*)m->loss_partial, (void*)m->losses, (void*)m->packed})
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* efvm_mlp_last_error(const efvm_mlp* m) { return m ? m->err.c_str() : "null network"; }

efvm_status efvm_mlp_load(efvm_mlp* m, const float* params, size_t n) {
    if (!m || !params) return EFVM_ERR_ARG;
    if (n != (size_t)mlp::NPARAM) return mfail(m, EFVM_ERR_ARG, "parameter count must be " + std::to_string(mlp::NPARAM));
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(params[i])) return mfail(m, EFVM_ERR_ARG, "non-finite parameter");
    MHIPCHK(m, hipMemcpyAsync(m->p, params, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipMemsetAsync(m->m, 0, sizeof(float) * n, m->stream));
    MHIPCHK(m, hipMemsetAsync(m->v, 0, sizeof(float) * n, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    m->packed_valid = 0;
    return mlp_repack(m);
}

efvm_status efvm_mlp_save(efvm_mlp* m, float* params, size_t n) {
    if (!m || !params) return EFVM_ERR_ARG;
    if (n != (size_t)mlp::NPARAM) return mfail(m, EFVM_ERR_ARG, "parameter count must be " + std::to_string(mlp::NPARAM));
    MHIPCHK(m, hipMemcpyAsync(params, m->p, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return EFVM_OK;
}

efvm_status efvm_mlp_clear_samples(efvm_mlp* m) {
    if (!m) return EFVM_ERR_ARG;
    m->n_data = 0;
    return EFVM_OK;
}

efvm_status efvm_mlp_num_samples(efvm_mlp* m, size_t* n) {
    if (!m || !n) return EFVM_ERR_ARG;
    *n = (size_t)m->n_data;
    return EFVM_OK;
}

efvm_status efvm_mlp_add_samples(efvm_mlp* m, const float* X, const float* Y, size_t n) {
    if (!m || (n && (!X || !Y))) return EFVM_ERR_ARG;
    efvm_status st = mlp_grow(m, m->n_data + (long)n);
    if (st != EFVM_OK) return st;
    MHIPCHK(m, hipMemcpyAsync(m->X + mlp::NIN * m->n_data, X, sizeof(float) * mlp::NIN * n, hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipMemcpyAsync(m->Y + mlp::NOUT * m->n_data, Y, sizeof(float) * mlp::NOUT * n, hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    m->n_data += (long)n;
    return EFVM_OK;
}

efvm_status efvm_mlp_get_samples(efvm_mlp* m, size_t first, size_t n, float* X, float* Y) {
    if (!m || !X || !Y) return EFVM_ERR_ARG;
    if ((long)(first + n) > m->n_data) return mfail(m, EFVM_ERR_ARG, "sample range out of bounds");
    MHIPCHK(m, hipMemcpyAsync(X, m->X + mlp::NIN * first, sizeof(float) * mlp::NIN * n, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipMemcpyAsync(Y, m->Y + mlp::NOUT * first, sizeof(float) * mlp::NOUT * n, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return EFVM_OK;
}

efvm_status efvm_mlp_add_snapshot(efvm_mlp* m, efvm_ctx* fine, int factor) {
    return efvm_mlp_add_snapshot_pair(m, fine, nullptr, factor);
}

efvm_status efvm_mlp_add_snapshot_pair(efvm_mlp* m, efvm_ctx* fine, efvm_ctx* coarse, int factor) {
    if (!m || !fine || factor < 1) return EFVM_ERR_ARG;
    const int nx = fine->grid.nx, ny = fine->grid.ny;
    if (!periodic_axis(fine, 0) || !periodic_axis(fine, 1))
        return mfail(m, EFVM_ERR_UNSUPPORTED, "training snapshots need a periodic grid");
    if (nx % factor || ny % factor || nx / factor < 5 || ny / factor < 5)
        return mfail(m, EFVM_ERR_ARG, "grid must divide by the factor into at least 5x5 coarse cells");
    if (fine->rk_stage != 0) return mfail(m, EFVM_ERR_ARG, "snapshot at a step boundary only");
    const long ncx = nx / factor, ncy = ny / factor, nc = ncx * ncy;
    if (coarse) {
        if (coarse->grid.nx != ncx || coarse->grid.ny != ncy || coarse->rk_stage != 0)
            return mfail(m, EFVM_ERR_ARG, "coarse context must be the fine grid divided by the factor, at a step boundary");
        MHIPCHK(m, hipStreamSynchronize(coarse->stream));
    }
    const long n_out = 4 * nc;
    efvm_status st = mlp_grow(m, m->n_data + n_out);
    if (st != EFVM_OK) return st;
    MHIPCHK(m, hipStreamSynchronize(fine->stream));
    double* tmp = nullptr;
    MHIPCHK(m, hipMalloc((void**)&tmp, sizeof(double) * 12 * nc));
    mlp::SampleArgs a;
    a.nxf = nx; a.nyf = ny; a.f = factor; a.nc_x = (int)ncx; a.nc_y = (int)ncy;
    a.gamma = fine->gas.gamma;
    for (int q = 0; q < 4; q++) a.U[q] = fine->P[fine->cur][q];
    a.esz = (int)fine->esz;
    for (int q = 0; q < 4; q++) a.Uc[q] = coarse ? coarse->P[coarse->cur][q] : nullptr;
    a.esz_c = coarse ? (int)coarse->esz : 0;
    a.cells = tmp;
    a.faces = tmp + 4 * nc;
    a.X = m->X + mlp::NIN * m->n_data;
    a.Y = m->Y + mlp::NOUT * m->n_data;
    hipError_t e = launch_mlp_samples(a, n_out, m->stream);
    hipError_t e2 = hipStreamSynchronize(m->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess || e2 != hipSuccess)
        return mfail(m, EFVM_ERR_HIP, std::string("sample kernels: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    m->n_data += n_out;
    return EFVM_OK;
}

efvm_status efvm_mlp_gradient(efvm_mlp* m, const float* X, const float* Y, size_t n, float* grad, double* loss) {
    if (!m || !X || !Y || !grad || n == 0) return EFVM_ERR_ARG;
    float *dX = nullptr, *dY = nullptr;
    MHIPCHK(m, hipMalloc((void**)&dX, sizeof(float) * mlp::NIN * n));
    MHIPCHK(m, hipMalloc((void**)&dY, sizeof(float) * mlp::NOUT * n));
    const long nblk = ((long)n + mlp::TB - 1) / mlp::TB;
    efvm_status st = mlp_reserve_partials(m, nblk);
    if (st == EFVM_OK && m->loss_cap < 1) {
        if (hipMalloc((void**)&m->losses, sizeof(double) * 1024) == hipSuccess) m->loss_cap = 1024;
        else st = mfail(m, EFVM_ERR_HIP, "hipMalloc losses");
    }
    if (st == EFVM_OK) {
        hipError_t e = hipMemcpyAsync(dX, X, sizeof(float) * mlp::NIN * n, hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dY, Y, sizeof(float) * mlp::NOUT * n, hipMemcpyHostToDevice, m->stream);
        mlp::TrainArgs a;
        a.params = m->p; a.X = dX; a.Y = dY; a.n_data = (long)n; a.n = (long)n; a.random = 0; a.seed = 0; a.step = 0;
        a.scale = (float)(2.0 / ((double)n * mlp::NOUT));
        a.partial = m->partial; a.loss_partial = m->loss_partial;
        if (e == hipSuccess) e = launch_mlp_train(a, (int)nblk, m->g, m->losses, m->stream);
        double l = 0.0;
        if (e == hipSuccess) e = hipMemcpyAsync(grad, m->g, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&l, m->losses, sizeof(double), hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) st = mfail(m, EFVM_ERR_HIP, std::string("training kernels: ") + hipGetErrorString(e));
        if (loss) *loss = l;
    }
    (void)hipFree(dX);
    (void)hipFree(dY);
    return st;
}

efvm_status efvm_mlp_train_step(efvm_mlp* m, int n_steps, int batch, float lr0, float lr1, unsigned long long seed,
                                int train_skip, double* losses) {
    if (!m || n_steps < 0 || batch < 1) return EFVM_ERR_ARG;
    if (m->n_data < 1) return mfail(m, EFVM_ERR_ARG, "no training samples");
    const long nblk = ((long)batch + mlp::TB - 1) / mlp::TB;
    efvm_status st = mlp_reserve_partials(m, nblk);
    if (st != EFVM_OK) return st;
    if (m->loss_cap < n_steps) {
        if (m->losses) (void)hipFree(m->losses);
        m->losses = nullptr; m->loss_cap = 0;
        MHIPCHK(m, hipMalloc((void**)&m->losses, sizeof(double) * std::max(n_steps, 1024)));
        m->loss_cap = std::max(n_steps, 1024);
    }
    for (int k = 0; k < n_steps; k++) {
        // cosine schedule from lr0 to lr1 over this call's steps
        const double f = n_steps > 1 ? (double)k / (double)(n_steps - 1) : 0.0;
        const float lr = (float)(lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(M_PI * f)));
        const long t = ++m->adam_t;
        mlp::TrainArgs a;
        a.params = m->p; a.X = m->X; a.Y = m->Y; a.n_data = m->n_data; a.n = batch; a.random = 1;
        a.seed = seed; a.step = (unsigned long long)t;
        a.scale = (float)(2.0 / ((double)batch * mlp::NOUT));
        a.partial = m->partial; a.loss_partial = m->loss_partial;
        MHIPCHK(m, launch_mlp_train(a, (int)nblk, m->g, m->losses + k, m->stream));
        MHIPCHK(m, launch_mlp_adam(m->p, m->g, m->m, m->v, lr, (int)std::min<long>(t, 1 << 30),
                                   train_skip ? mlp::NPARAM : mlp::OFF_L, m->stream));
    }
    if (losses && n_steps)
        MHIPCHK(m, hipMemcpyAsync(losses, m->losses, sizeof(double) * n_steps, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    m->packed_valid = 0;
    return mlp_repack(m);
}

efvm_status efvm_mlp_fit_linear(efvm_mlp* m, size_t stride, double ridge, int constrained) {
    if (!m || stride < 1 || ridge < 0.0 || constrained < 0 || constrained > 1) return EFVM_ERR_ARG;
    if (m->n_data < 1) return mfail(m, EFVM_ERR_ARG, "no training samples");
    constexpr int NA = mlp::NIN + 1, NG = mlp::NOUT * NA * (NA + 1);
    const int blocks = 512;
    double* part = nullptr;
    MHIPCHK(m, hipMalloc((void**)&part, sizeof(double) * NG * blocks));
    std::vector<double> h((size_t)NG * blocks);
    hipError_t e = launch_mlp_gram(m->X, m->Y, m->n_data, (long)stride, part, blocks, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return mfail(m, EFVM_ERR_HIP, std::string("normal-equation kernel: ") + hipGetErrorString(e));
    std::vector<float> p(mlp::NPARAM);
    MHIPCHK(m, hipMemcpyAsync(p.data(), m->p, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    for (int v = 0; v < mlp::NOUT; v++) {
        // sum the block partials, then solve (G + ridge tr(G)/NA I) c = b by Gaussian elimination with pivoting
        double A[NA][NA + 1];
        for (int i = 0; i < NA; i++)
            for (int j = 0; j <= NA; j++) {
                double s = 0.0;
                for (int b = 0; b < blocks; b++) s += h[(size_t)b * NG + (size_t)v * NA * (NA + 1) + i * (NA + 1) + j];
                A[i][j] = s;
            }
        double tr = 0.0;
        for (int i = 0; i < NA; i++) tr += A[i][i];
        if (!(tr > 0.0)) return mfail(m, EFVM_ERR_NONPHYSICAL, "degenerate normal equations (no weighted samples)");
        if (constrained) {
            // Offsets exact for polynomial profiles up to degree 3 on the own variable: the
            // coefficients of x[-2], x[-1], x[+1], x[+2] are the fifth-order face value's
            // (2, -13, 27, -3) / 60, less the central base (0, -1/4, +1/4, 0), plus t (1, -4, -4, 1),
            // the one direction that keeps exactness; every other coefficient is 0.  t is the
            // weighted least-squares optimum along that line: t = d.(b - G c0) / (d.G d).
            double c0[NA] = {}, d[NA] = {};
            const int kk[4] = {0, 1, 3, 4};
            const double lin5[4] = {2.0 / 60.0, -13.0 / 60.0 + 0.25, 27.0 / 60.0 - 0.25, -3.0 / 60.0};
            const double dir[4] = {1.0, -4.0, -4.0, 1.0};
            for (int q = 0; q < 4; q++) { c0[4 * kk[q] + v] = lin5[q]; d[4 * kk[q] + v] = dir[q]; }
            double num = 0.0, den = 0.0;
            for (int i = 0; i < NA; i++) {
                double Gc = 0.0, Gd = 0.0;
                for (int j = 0; j < NA; j++) { Gc += A[i][j] * c0[j]; Gd += A[i][j] * d[j]; }
                num += d[i] * (A[i][NA] - Gc);
                den += d[i] * Gd;
            }
            const double t = den > 0.0 ? num / (den + ridge * tr) : 0.0;
            for (int i = 0; i < mlp::NIN; i++) p[mlp::OFF_L + v * mlp::NIN + i] = (float)(c0[i] + t * d[i]);
            p[mlp::OFF_BL + v] = 0.f;
            continue;
        }
        for (int i = 0; i < NA; i++) A[i][i] += ridge * tr / NA;
        for (int c = 0; c < NA; c++) {
            int piv = c;
            for (int r = c + 1; r < NA; r++) if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
            if (std::fabs(A[piv][c]) < 1e-300) return mfail(m, EFVM_ERR_NONPHYSICAL, "singular normal equations");
            if (piv != c) for (int j = 0; j <= NA; j++) std::swap(A[c][j], A[piv][j]);
            for (int r = 0; r < NA; r++) {
                if (r == c) continue;
                const double f = A[r][c] / A[c][c];
                for (int j = c; j <= NA; j++) A[r][j] -= f * A[c][j];
            }
        }
        for (int i = 0; i < mlp::NIN; i++) p[mlp::OFF_L + v * mlp::NIN + i] = (float)(A[i][NA] / A[i][i]);
        p[mlp::OFF_BL + v] = (float)(A[mlp::NIN][NA] / A[mlp::NIN][mlp::NIN]);
    }
    MHIPCHK(m, hipMemcpyAsync(m->p, p.data(), sizeof(float) * mlp::NPARAM, hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    m->packed_valid = 0;
    return mlp_repack(m);
}

efvm_status efvm_mlp_infer(efvm_mlp* m, const float* X, size_t n, float* Y) {
    if (!m || !X || !Y || n == 0) return EFVM_ERR_ARG;
    efvm_status st = mlp_repack(m);
    if (st != EFVM_OK) return st;
    float *dX = nullptr, *dY = nullptr;
    MHIPCHK(m, hipMalloc((void**)&dX, sizeof(float) * mlp::NIN * n));
    MHIPCHK(m, hipMalloc((void**)&dY, sizeof(float) * mlp::NOUT * n));
    hipError_t e = hipMemcpyAsync(dX, X, sizeof(float) * mlp::NIN * n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = launch_mlp_infer(m->packed, dX, (long)n, dY, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(Y, dY, sizeof(float) * mlp::NOUT * n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (e != hipSuccess) return mfail(m, EFVM_ERR_HIP, std::string("inference kernel: ") + hipGetErrorString(e));
    return EFVM_OK;
}

efvm_status efvm_set_learned(efvm_ctx* c, efvm_mlp* m) {
    if (!c) return EFVM_ERR_ARG;
    drop_graph(c);
    if (!m) { c->learned = 0; return EFVM_OK; }
    if (c->esz != 4) return fail(c, EFVM_ERR_UNSUPPORTED, "learned reconstruction runs in FP32");
    if (!periodic_axis(c, 0) || !periodic_axis(c, 1)) return fail(c, EFVM_ERR_UNSUPPORTED, "learned reconstruction needs a periodic grid");
    if (c->sch.viscous || c->n_cyl > 0 || c->body_force || c->sponge || c->parabolic)
        return fail(c, EFVM_ERR_UNSUPPORTED, "learned reconstruction covers the inviscid periodic path only");
    if (c->grid.nx < 5 || c->grid.ny < 5) return fail(c, EFVM_ERR_CONFIG, "learned reconstruction needs at least 5x5 cells");
    efvm_status st = mlp_repack(m);
    if (st != EFVM_OK) return fail(c, st, m->err);
    if (!c->faces) HIPCHK(c, hipMalloc((void**)&c->faces, sizeof(float) * 16 * c->ncell));
    if (!c->wpack) HIPCHK(c, hipMalloc((void**)&c->wpack, sizeof(mlp::Packed)));
    if (!c->lstats) {
        HIPCHK(c, hipMalloc((void**)&c->lstats, 3 * mlp::STAT_SLOTS * sizeof(unsigned long long)));
        HIPCHK(c, hipMemsetAsync(c->lstats, 0, 3 * mlp::STAT_SLOTS * sizeof(unsigned long long), c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->wpack, m->packed, sizeof(mlp::Packed), hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(c, hipStreamSynchronize(m->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->learned = 1;
    c->learned_own4 = m->own4;
    std::memcpy(c->learned_taps, m->taps, sizeof(c->learned_taps));
    return EFVM_OK;
}

efvm_status efvm_set_learned_bypass(efvm_ctx* c, int bypass) {
    if (!c || (bypass != 0 && bypass != 1)) return EFVM_ERR_ARG;
    if (bypass && 2.0 * (double)c->ncell >= 4294967295.0)
        return fail(c, EFVM_ERR_UNSUPPORTED, "bypass face list indexes faces in 32 bits: grid too large");
    drop_graph(c);
    if (bypass && !c->dlist) {
        HIPCHK(c, hipMalloc((void**)&c->dlist, sizeof(unsigned) * (2 * c->ncell + 4)));
        HIPCHK(c, hipMemsetAsync(c->dlist + 2 * c->ncell, 0, 4 * sizeof(unsigned), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->learned_bypass = bypass;
    return EFVM_OK;
}

efvm_status efvm_set_learned_fused(efvm_ctx* c, int allow) {
    if (!c || (allow != 0 && allow != 1)) return EFVM_ERR_ARG;
    c->learned_fused = allow;
    drop_graph(c);
    return EFVM_OK;
}

efvm_status efvm_get_learned_fused(const efvm_ctx* c, int* allow, int* fits) {
    if (!c) return EFVM_ERR_ARG;
    if (allow) *allow = c->learned_fused;
    if (fits) *fits = learned_fused_fits(c) ? 1 : 0;
    return EFVM_OK;
}

efvm_status efvm_mlp_infer_stats(efvm_ctx* c, int reset, unsigned long long out[3]) {
    if (!c || !out) return EFVM_ERR_ARG;
    out[0] = out[1] = out[2] = 0;
    if (!c->lstats) return EFVM_OK;
    std::vector<unsigned long long> h(3 * mlp::STAT_SLOTS);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->lstats, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    if (reset) HIPCHK(c, hipMemsetAsync(c->lstats, 0, h.size() * sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int w = 0; w < mlp::STAT_SLOTS; w++)
        for (int k = 0; k < 3; k++) out[k] += h[3 * w + k];
    return EFVM_OK;
}

efvm_status efvm_learned_faces(efvm_ctx* c, float* out, size_t n) {
    if (!c || !out) return EFVM_ERR_ARG;
    if (!c->learned) return fail(c, EFVM_ERR_ARG, "no learned reconstruction attached");
    if ((long)n != 16 * c->ncell) return fail(c, EFVM_ERR_ARG, "face buffer must hold 16 * nx * ny floats");
    if (c->rk_stage != 0) return fail(c, EFVM_ERR_ARG, "face states at a step boundary only");
    mlp::FaceArgs fa;
    fa.nx = c->grid.nx; fa.ny = c->grid.ny;
    fa.n_chunks = (c->grid.nx + 63) / 64;
    fa.n_tasks = 2 * c->grid.ny * fa.n_chunks;
    fa.gamma = (float)c->gas.gamma;
    for (int q = 0; q < 4; q++) fa.U[q] = (const float*)c->P[c->cur][q];
    fa.F = c->faces;
    fa.w = c->wpack;
    fa.stats = nullptr;
    fa.bypass = c->learned_bypass;
    fa.out_mode = 0;
    fa.flux = c->sch.flux;
    fa.list = nullptr;
    fa.n_list = nullptr;
    fa.own4 = c->learned_own4;
    fa.n_lam_slots = 0;
    std::memcpy(fa.taps, c->learned_taps, sizeof(fa.taps));
    HIPCHK(c, launch_mlp_faces(fa, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, c->faces, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// State I/O: CFDRST01 checkpoints (examples/restart_determinism/main.rs:177-226)
// and VTK legacy RECTILINEAR_GRID output (crates/cfd_output/src/lib.rs:36-113).
// ---------------------------------------------------------------------------
namespace {

// the reference's raw index of interior cell (i, j): ng = 2 ghost layers on every axis of
// the one-cell-thick mesh, x fastest (nx + 4) (ny + 4) (1 + 4) cells
inline long raw_index(int nx, int ny, int i, int j) {
    const long NX = nx + 4, NY = ny + 4;
    return (long)(i + 2) + NX * ((long)(j + 2) + NY * 2);
}

efvm_status download_set(efvm_ctx* c, int set, double* cons5) {
    const long n = c->ncell;
    std::vector<char> h[4];
    for (int m = 0; m < 4; m++) {
        h[m].resize(c->esz * n);
        HIPCHK(c, hipMemcpyAsync(h[m].data(), c->P[set][m], c->esz * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int comp[4] = {0, 1, 2, 4};
    for (long q = 0; q < n; q++) {
        cons5[5 * q + 3] = 0.0;
        for (int m = 0; m < 4; m++)
            cons5[5 * q + comp[m]] = c->esz == 8 ? reinterpret_cast<double*>(h[m].data())[q]
                                                 : (double)reinterpret_cast<float*>(h[m].data())[q];
    }
    return EFVM_OK;
}

template <class T> void put(std::vector<unsigned char>& b, T v) {
    unsigned char x[sizeof(T)];
    std::memcpy(x, &v, sizeof(T));   // little-endian host (x86-64), as the format is
    b.insert(b.end(), x, x + sizeof(T));
}

// VTK-safe scientific notation: 10 digits, signed exponent of at least two digits,
// non-finite values as zero (crates/cfd_output/src/lib.rs:15-33)
void fmt_e(FILE* f, double x) {
    if (!std::isfinite(x)) x = 0.0;
    std::fprintf(f, "%.10e", x);
}

}  // namespace

namespace efvm {
// write_vtk_to (crates/cfd_output/src/lib.rs:36-113): RECTILINEAR_GRID over nn[0] x nn[1] x nn[2]
// cells (x fastest), face coordinates from the centroids, density, pressure, temperature,
// velocity and Mach in %.10e with non-finite values written as 0.
efvm_status vtk_write_cells(const int nn[3], const double lo[3], const double h[3], double gm, double R,
                            const double* cons5, const char* path, const double* const* edges) {
    FILE* f = std::fopen(path, "w");
    if (!f) return EFVM_ERR_ARG;
    std::fprintf(f, "# vtk DataFile Version 3.0\ndev_field_efvm output\nASCII\nDATASET RECTILINEAR_GRID\n");
    std::fprintf(f, "DIMENSIONS %d %d %d\n", nn[0] + 1, nn[1] + 1, nn[2] + 1);
    // faces = centroid - half spacing, plus the last centroid + half spacing
    const char* label[3] = {"X", "Y", "Z"};
    for (int ax = 0; ax < 3; ax++) {
        // a stretched axis: cell m on [e_m, e_{m+1}], centroid their midpoint, spacing their difference
        const double* e = edges ? edges[ax] : nullptr;
        auto cen = [&](int m) { return e ? 0.5 * (e[m] + e[m + 1]) : lo[ax] + ((double)m + 0.5) * h[ax]; };
        auto wid = [&](int m) { return e ? e[m + 1] - e[m] : h[ax]; };
        std::fprintf(f, "%s_COORDINATES %d double\n", label[ax], nn[ax] + 1);
        for (int m = 0; m < nn[ax]; m++) {
            fmt_e(f, cen(m) - 0.5 * wid(m));
            std::fputc('\n', f);
        }
        fmt_e(f, cen(nn[ax] - 1) + 0.5 * wid(nn[ax] - 1));
        std::fputc('\n', f);
    }
    const long nc = (long)nn[0] * nn[1] * nn[2];
    std::fprintf(f, "CELL_DATA %ld\n", nc);
    auto cell = [&](long q) { return C5<double>{cons5[5 * q], cons5[5 * q + 1], cons5[5 * q + 2], cons5[5 * q + 3], cons5[5 * q + 4]}; };
    auto scalar = [&](const char* name, auto fn) {
        std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", name);
        for (long q = 0; q < nc; q++) {
            fmt_e(f, fn(cell(q)));
            std::fputc('\n', f);
        }
    };
    scalar("density", [](const C5<double>& u) { return u.r; });
    scalar("pressure", [&](const C5<double>& u) { return ref::pressure(u, gm); });
    scalar("temperature", [&](const C5<double>& u) { return ref::cons_to_prim(u, gm, R).t; });
    std::fprintf(f, "VECTORS velocity double\n");
    for (long q = 0; q < nc; q++) {
        double v[3];
        ref::velocity(cell(q), v);
        fmt_e(f, v[0]); std::fputc(' ', f); fmt_e(f, v[1]); std::fputc(' ', f); fmt_e(f, v[2]); std::fputc('\n', f);
    }
    scalar("mach", [&](const C5<double>& u) {
        double v[3];
        ref::velocity(u, v);
        return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / ref::sound(u, gm);
    });
    return std::fclose(f) == 0 ? EFVM_OK : EFVM_ERR_ARG;
}
}  // namespace efvm


extern "C" {

efvm_status efvm_checkpoint_save(efvm_ctx* c, const char* path) {
    if (!c || !path) return EFVM_ERR_ARG;
    const int nx = c->grid.nx, ny = c->grid.ny;
    const long n = c->ncell, nraw = (long)(nx + 4) * (ny + 4) * 5;
    std::vector<double> u(5 * n), u0(5 * n, 0.0);
    efvm_status st = efvm_download_state(c, u.data(), n);          // the current substage's state
    if (st != EFVM_OK) return st;
    // the reference's u0 (save_state at the first substage, crates/cfd_solver/src/integrate.rs:76-78):
    // mid-step U^n; at a step boundary the state before the completed step, U^{n-1}, which the
    // plane-set rotation keeps (zeros before the first step, as a new CfdState holds)
    const bool rk3 = c->sch.integrator == EFVM_INTEGRATOR_RK3;   // Euler never saves: u0 stays zero
    const int u0_set = !rk3 ? -1 : (c->rk_stage != 0 ? c->cur : (c->have_prev ? c->prev : -1));
    if (u0_set >= 0 && (st = download_set(c, u0_set, u0.data())) != EFVM_OK) return st;
    double dt = 0.0, time = 0.0;
    long step = 0;
    int stage = 0;
    if ((st = efvm_get_clock(c, &dt, &time, &step, &stage)) != EFVM_OK) return st;
    std::vector<unsigned char> b;
    b.reserve(8 + 25 + 3 * (8 + 40 * nraw));
    const char magic[8] = {'C', 'F', 'D', 'R', 'S', 'T', '0', '1'};
    b.insert(b.end(), magic, magic + 8);
    put(b, dt);
    put(b, time);
    put(b, (unsigned long long)step);
    put(b, (unsigned char)stage);
    // u, rhs, u0 over the raw padded array.  The device keeps neither ghost cells (the loaders
    // form them from the interior) nor an rhs column (the divergence lives in registers): ghost
    // cells and the rhs column are written as zeros.  Interior u and u0 are the reference's.
    const double* cols[3] = {u.data(), nullptr, u0.data()};
    for (int k = 0; k < 3; k++) {
        put(b, (unsigned long long)nraw);
        std::vector<double> raw(5 * nraw, 0.0);
        if (cols[k])
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++)
                    std::memcpy(&raw[5 * raw_index(nx, ny, i, j)], cols[k] + 5 * ((long)j * nx + i), 5 * sizeof(double));
        const unsigned char* p = reinterpret_cast<const unsigned char*>(raw.data());
        b.insert(b.end(), p, p + raw.size() * sizeof(double));
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, EFVM_ERR_ARG, std::string("checkpoint: cannot open ") + path);
    const size_t w = std::fwrite(b.data(), 1, b.size(), f);
    const int cl = std::fclose(f);
    if (w != b.size() || cl != 0) return fail(c, EFVM_ERR_ARG, std::string("checkpoint: short write to ") + path);
    return EFVM_OK;
}

efvm_status efvm_checkpoint_load(efvm_ctx* c, const char* path) {
    if (!c || !path) return EFVM_ERR_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, EFVM_ERR_ARG, std::string("checkpoint: cannot open ") + path);
    std::vector<unsigned char> b;
    unsigned char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + r);
    std::fclose(f);
    size_t pos = 0;
    auto take = [&](void* dst, size_t k) {
        if (pos + k > b.size()) return false;
        std::memcpy(dst, b.data() + pos, k);
        pos += k;
        return true;
    };
    char magic[8];
    double dt, time;
    unsigned long long step;
    unsigned char stage;
    if (!take(magic, 8) || std::memcmp(magic, "CFDRST01", 8) != 0)
        return fail(c, EFVM_ERR_CONFIG, "checkpoint: not a CFDRST01 file");
    if (!take(&dt, 8) || !take(&time, 8) || !take(&step, 8) || !take(&stage, 1))
        return fail(c, EFVM_ERR_CONFIG, "checkpoint: truncated header");
    if (stage > 2) return fail(c, EFVM_ERR_CONFIG, "checkpoint: bad rk stage");
    if (stage != 0)
        return fail(c, EFVM_ERR_UNSUPPORTED, "checkpoint: restarts are taken at a step boundary (rk stage 0)");
    const int nx = c->grid.nx, ny = c->grid.ny;
    const long n = c->ncell, nraw = (long)(nx + 4) * (ny + 4) * 5;
    std::vector<double> u(5 * n), u0(5 * n);
    for (int k = 0; k < 3; k++) {
        unsigned long long len;
        if (!take(&len, 8)) return fail(c, EFVM_ERR_CONFIG, "checkpoint: truncated column");
        if ((long)len != nraw) return fail(c, EFVM_ERR_CONFIG, "checkpoint: cell count does not match this grid");
        if (pos + 40 * (size_t)nraw > b.size()) return fail(c, EFVM_ERR_CONFIG, "checkpoint: truncated column");
        double* dst = k == 0 ? u.data() : (k == 2 ? u0.data() : nullptr);
        if (dst)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++)
                    std::memcpy(dst + 5 * ((long)j * nx + i), b.data() + pos + 40 * (size_t)raw_index(nx, ny, i, j), 40);
        pos += 40 * (size_t)nraw;
    }
    if (pos != b.size()) return fail(c, EFVM_ERR_CONFIG, "checkpoint: trailing bytes");
    for (long q = 0; q < n; q++)
        if (u0[5 * q + 3] != 0.0) return fail(c, EFVM_ERR_UNSUPPORTED, "checkpoint: rho_w of u0 must be 0 on the 2-D GPU path");
    efvm_status st = efvm_upload_state(c, u.data(), n);      // rk stage 0; the next step recomputes the CFL
    if (st != EFVM_OK) return st;
    // the file's u0 (the state before its last step) goes to the prev set: a checkpoint written
    // next carries it unchanged; the first step overwrites it
    if ((st = upload_set(c, c->prev, u0.data(), n)) != EFVM_OK) return st;
    c->have_prev = 1;
    drop_graph(c);
    HIPCHK(c, launch_clock_set(c->clock, dt, time, (long long)step, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm_write_vtk(const efvm_grid* g, const efvm_gas* gas, const double* cons5, size_t n_cells, const char* path) {
    if (!g || !gas || !cons5 || !path || g->nx < 1 || g->ny < 1) return EFVM_ERR_ARG;
    if ((long)n_cells != (long)g->nx * g->ny) return EFVM_ERR_ARG;
    const int nn[3] = {g->nx, g->ny, 1};
    const double h[3] = {(g->hi_x - g->lo_x) / (double)g->nx, (g->hi_y - g->lo_y) / (double)g->ny, g->dz};
    const double lo[3] = {g->lo_x, g->lo_y, 0.0};
    return efvm::vtk_write_cells(nn, lo, h, gas->gamma, gas->r_gas, cons5, path, nullptr);
}

}  // extern "C"
