// efvm_api3d.cpp -- C ABI of the 3-D structured-mesh context (efvm3_*).
//
// The host side of one App over a UniformMesh with nz cells along z: it owns the
// device planes (three raw padded state sets, rotated through the RK substages),
// the clock, the immersed-body tables, and enqueues the phase kernels of
// efvm_stage3d.hip on its stream.  Compiled with -ffp-contract=off: the geometry
// built here (viscous weights, image points, markers) is FP64 in the reference's
// operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/efvm.h"
#include "efvm_stage3d.h"

namespace efvm {
efvm_status vtk_write_cells(const int nn[3], const double lo[3], const double h[3], double gm, double R,
                            const double* cons5, const char* path, const double* const* edges);
}

using namespace efvm;

static constexpr int NG3 = 2;

struct efvm_ctx3 {
    efvm_grid3 grid{};
    efvm_gas gas{};
    efvm_scheme sch{};
    efvm_bc bc[6]{};
    double h[3] = {0, 0, 0};
    int NX = 0, NY = 0, NZ = 0;
    long nraw = 0, ncell = 0;
    hipStream_t stream = nullptr;
    size_t esz = 8;
    void* P[3][5] = {};      // raw padded state sets A, B, C
    int cur = 0;             // set holding U^n
    int prev = 1;            // set holding U^{n-1} (U1 during a step); the third set holds U2
    int have_prev = 0;       // prev holds the state before the last completed step (or a loaded u0)
    int rhs_state = 0;       // the checkpoint's rhs column: 0 zeros (no substage since an upload),
                             // 1 recomputed from the last substage's input, 2 the loaded column
    std::vector<double> loaded_rhs;
    DevClock* clock = nullptr;
    double* part = nullptr;  // CFL per-block maxima, then sums partials (5 per block)
    double* out = nullptr;   // marker force (6) and sums (5)
    double* yw = nullptr;
    double* zw = nullptr;
    std::vector<double> edges[3];  // stretched y / z: the n+1 cell edges (empty: uniform spacing h)
    double* width_d[3] = {};      // device: interior cell widths of a stretched axis
    double* cen_d[3] = {};        // device: interior cell centroids of a stretched axis
    uint8_t* solid = nullptr;     // raw-indexed
    long* ghost_idx = nullptr;
    double* ghost_img = nullptr;
    int n_ghost = 0;
    double* markers = nullptr;
    double* mforce = nullptr;     // per-marker force contributions (6 per marker)
    int n_mark = 0;
    double* stagnant = nullptr;
    double* bc_tab[6] = {};       // FromOutside ghost tables (6 doubles per face), allocated on first use
    int ibm_flipped = 1;
    int body_kind = 0;            // 0 none, 1 cylinders, 2 spheres
    int n_body = 0;
    double body[8][4] = {};       // cx, cy, cz, r (cylinders: cz unused)
    int sphere_markers = 2000;
    double freestream[5] = {};
    int rk_stage = 0;
    int parity = 0;
    int sponge = 0;
    Sponge3 sp{};
    int body_force = 0;
    double gx = 0.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_cur = 0, graph_parity = 0;   // plane set and CFL slot the captured steps start from
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int timed = 0;
    std::string err;
};

namespace {

efvm_status fail3(efvm_ctx3* c, efvm_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}
#define HIPCHK3(ctx, expr)                                                                               \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return fail3(ctx, EFVM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

bool periodic3(const efvm_ctx3* c, int axis) { return c->bc[2 * axis].kind == EFVM_BC_PERIODIC; }
// Cell centroid along axis d at raw position `raw`.  A stretched axis (UniformMeshConfig y_edges /
// z_edges) has cell i on [e_i, e_{i+1}] with centroid (e_i + e_{i+1}) / 2; its ghost centroids
// are the periodic images of the opposite cells on a periodic axis, else the mirror images of the
// interior cells across the boundary edge.
double centroid3(const efvm_ctx3* c, int d, int raw) {
    const std::vector<double>& e = c->edges[d];
    if (e.empty()) return c->grid.lo[d] + ((double)(raw - NG3) + 0.5) * c->h[d];
    const int n = (int)e.size() - 1, i = raw - NG3;
    auto cc = [&](int q) { return 0.5 * (e[q] + e[q + 1]); };
    if (i >= 0 && i < n) return cc(i);
    const int g = i < 0 ? -i : i - n + 1;   // ghost layer 1 or 2
    const double len = e[n] - e[0];
    if (periodic3(c, d)) return i < 0 ? cc(n - g) - len : cc(g - 1) + len;
    return i < 0 ? 2.0 * e[0] - cc(g - 1) : 2.0 * e[n] - cc(n - g);
}
long raw_index(const efvm_ctx3* c, int i, int j, int k) { return i + (long)c->NX * (j + (long)c->NY * k); }

Args3 args3(const efvm_ctx3* c) {
    Args3 a;
    std::memset(&a, 0, sizeof(a));
    a.nx = c->grid.nx; a.ny = c->grid.ny; a.nz = c->grid.nz;
    a.NX = c->NX; a.NY = c->NY; a.NZ = c->NZ;
    a.nraw = c->nraw; a.ncell = c->ncell;
    for (int d = 0; d < 3; d++) { a.lo[d] = c->grid.lo[d]; a.h[d] = c->h[d]; a.per[d] = periodic3(c, d); }
    const double vol = c->h[0] * c->h[1] * c->h[2];
    a.aov[0] = (c->h[1] * c->h[2]) / vol;
    a.aov[1] = (c->h[0] * c->h[2]) / vol;
    a.aov[2] = (c->h[0] * c->h[1]) / vol;
    a.seam_mode = c->sch.seam_mode;
    for (int s = 0; s < 6; s++) {
        a.bc_kind[s] = c->bc[s].kind;
        for (int q = 0; q < 5; q++) a.bc_par[s][q] = c->bc[s].par[q];
        if (c->bc[s].kind == EFVM_BC_SUPERSONIC_INFLOW) {
            const double* p = c->bc[s].par;
            P6<double> pv = {p[0], p[1], p[2], p[3], p[4], p[4] / (p[0] * 287.058)};
            C5<double> o = ref::prim_to_cons(pv, c->gas.gamma);
            a.bc_cons[s][0] = o.r; a.bc_cons[s][1] = o.ru; a.bc_cons[s][2] = o.rv; a.bc_cons[s][3] = o.rw; a.bc_cons[s][4] = o.re;
        }
    }
    a.gamma = c->gas.gamma; a.r_gas = c->gas.r_gas;
    a.flux = c->sch.flux; a.muscl = c->sch.recon == EFVM_RECON_MUSCL; a.limiter = c->sch.limiter;
    a.viscous = c->sch.viscous; a.visc_const = c->gas.viscosity == EFVM_VISC_CONSTANT; a.mu = c->gas.mu;
    a.dxi2 = 1.0 / (c->h[0] * c->h[0]);
    a.yw = c->yw; a.zw = c->zw;
    for (int d = 0; d < 3; d++) { a.spacing[d] = c->width_d[d]; a.cen[d] = c->cen_d[d]; }
    a.body_force = c->body_force; a.gx = c->gx;
    a.solid = c->n_body > 0 ? c->solid : nullptr;
    for (int s = 0; s < 6; s++) a.tab[s] = c->bc[s].kind == EFVM_BC_TABLE ? c->bc_tab[s] : nullptr;
    a.euler = c->sch.integrator == EFVM_INTEGRATOR_EULER;
    a.cfl = c->sch.cfl; a.fixed_dt = c->sch.fixed_dt;
    a.clock = c->clock;
    return a;
}

void planes3(Args3& a, void* const* in, void* const* u0, void* const* out) {
    for (int m = 0; m < 5; m++) { a.uin[m] = in[m]; a.u0[m] = u0[m]; a.uout[m] = out[m]; }
}

// the set holding the state of the current substage: A at a step boundary, else B / C
int current_set(const efvm_ctx3* c) {
    if (c->sch.integrator == EFVM_INTEGRATOR_RK3 && c->rk_stage == 1) return c->prev;
    if (c->sch.integrator == EFVM_INTEGRATOR_RK3 && c->rk_stage == 2) return 3 - c->cur - c->prev;
    return c->cur;
}

void drop_graph3(efvm_ctx3* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr; c->graph = nullptr;
}

void free_ibm3(efvm_ctx3* c) {
    if (c->solid) (void)hipFree(c->solid);
    if (c->ghost_idx) (void)hipFree(c->ghost_idx);
    if (c->ghost_img) (void)hipFree(c->ghost_img);
    if (c->markers) (void)hipFree(c->markers);
    if (c->mforce) (void)hipFree(c->mforce);
    if (c->stagnant) (void)hipFree(c->stagnant);
    c->solid = nullptr; c->ghost_idx = nullptr; c->ghost_img = nullptr; c->markers = nullptr; c->stagnant = nullptr;
    c->mforce = nullptr;
    c->n_ghost = 0; c->n_mark = 0;
}

#define LAUNCH3(ctx, expr, what)                                                                         \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return fail3(ctx, EFVM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(_e)); \
    } while (0)

// One app.run(): Setup (CFL on the first substage, mask refresh) -> Halo -> Boundary ->
// PreFlux (mirror) -> Flux .. Update (fused) -> PostUpdate (sponge).  Enqueue only.
efvm_status run3(efvm_ctx3* c) {
    const bool rk3 = c->sch.integrator == EFVM_INTEGRATOR_RK3;
    // roles: cur = U^n, prev = U^{n-1} (receives U1, then U^{n+1}), the third set U2; after the
    // step cur and prev swap, so U^n survives the step (the reference's u0 afterwards)
    void* const* A = c->P[c->cur];
    void* const* B = c->P[c->prev];
    void* const* Cc = c->P[3 - c->cur - c->prev];
    const int esz = (int)c->esz;
    Args3 a = args3(c);
    a.parity = c->parity;
    a.stage = rk3 ? c->rk_stage : 0;
    void* const* in = A;
    void* const* out = B;
    if (rk3 && c->rk_stage == 1) { in = B; out = Cc; }
    if (rk3 && c->rk_stage == 2) { in = Cc; out = B; }
    planes3(a, in, A, out);
    if (c->rk_stage == 0 && c->sch.fixed_dt <= 0.0)
        LAUNCH3(c, launch3_cfl(a, esz, c->part, &c->clock->lam[c->parity], c->stream), "cfl");
    if (c->n_body > 0 && !c->ibm_flipped) {
        LAUNCH3(c, launch3_flip(a, esz, c->stagnant, c->stream), "mask refresh");
        c->ibm_flipped = 1;
    }
    LAUNCH3(c, launch3_fill(a, esz, c->stream), "halo / boundary");
    if (c->n_body > 0 && c->n_ghost > 0)
        LAUNCH3(c, launch3_mirror(a, esz, c->n_ghost, c->ghost_idx, c->ghost_img, c->stream), "ibm mirror");
    LAUNCH3(c, launch3_stage(a, esz, c->stream), "stage");
    c->rhs_state = 1;
    const bool last = !rk3 || c->rk_stage == 2;
    if (last) {
        std::swap(c->cur, c->prev);
        c->have_prev = 1;
        if (c->sponge) {
            Args3 b = args3(c);
            planes3(b, c->P[c->cur], c->P[c->cur], c->P[c->cur]);
            LAUNCH3(c, launch3_sponge(b, esz, c->sp, c->stream), "sponge");
        }
        c->parity ^= 1;
        c->rk_stage = 0;
    } else {
        c->rk_stage++;
    }
    return EFVM_OK;
}

efvm_status alloc3(efvm_ctx3* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail3(c, EFVM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return EFVM_OK;
}

// viscous weights from cell-centre distances along one axis (crates/cfd_solver/src/lib.rs:483-501)
std::vector<double> axis_weights(const efvm_ctx3* c, int d, int n) {
    std::vector<double> w(3 * (size_t)n);
    for (int q = 0; q < n; q++) {
        const int r = q + NG3;
        const double hm = centroid3(c, d, r) - centroid3(c, d, r - 1), hp = centroid3(c, d, r + 1) - centroid3(c, d, r);
        w[q] = 1.0 / hm; w[n + q] = 1.0 / hp; w[2 * (size_t)n + q] = 2.0 / (hm + hp);
    }
    return w;
}

// The y and z viscous weight tables of the current geometry and boundaries.
efvm_status upload_weights3(efvm_ctx3* c) {
    const std::vector<double> yw = axis_weights(c, 1, c->grid.ny), zw = axis_weights(c, 2, c->grid.nz);
    HIPCHK3(c, hipMemcpyAsync(c->yw, yw.data(), sizeof(double) * yw.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK3(c, hipMemcpyAsync(c->zw, zw.data(), sizeof(double) * zw.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

// Signed distance and outward normal of body b at point p (sphere.rs:43-59; lib.rs cylinder).
double body_sd(const efvm_ctx3* c, int b, const double* p) {
    const double* B = c->body[b];
    const double dx = p[0] - B[0], dy = p[1] - B[1];
    if (c->body_kind == 2) {
        const double dz = p[2] - B[2];
        return std::sqrt(dx * dx + dy * dy + dz * dz) - B[3];
    }
    return std::sqrt(dx * dx + dy * dy) - B[3];
}
void body_normal(const efvm_ctx3* c, int b, const double* p, double* n) {
    const double* B = c->body[b];
    const double dx = p[0] - B[0], dy = p[1] - B[1];
    const double dz = c->body_kind == 2 ? p[2] - B[2] : 0.0;
    const double r = c->body_kind == 2 ? std::sqrt(dx * dx + dy * dy + dz * dz) : std::sqrt(dx * dx + dy * dy);
    if (r < 1e-12) { n[0] = 1.0; n[1] = 0.0; n[2] = 0.0; }
    else { n[0] = dx / r; n[1] = dy / r; n[2] = c->body_kind == 2 ? dz / r : 0.0; }
}

// Mask, ghost list and markers on the host in FP64, as the reference's Setup systems
// build them (refresh_*_mask, compute_*_ghosts, compute_*_markers); static bodies need
// them once.  The cells inside flip to the stagnant state on the next substage.
efvm_status build_ibm3(efvm_ctx3* c) {
    drop_graph3(c);
    free_ibm3(c);
    const int nx = c->grid.nx, ny = c->grid.ny, nz = c->grid.nz;
    std::vector<uint8_t> solid(c->nraw, 0);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double cc[3] = {centroid3(c, 0, i + NG3), centroid3(c, 1, j + NG3), centroid3(c, 2, k + NG3)};
                for (int b = 0; b < c->n_body; b++)
                    if (body_sd(c, b, cc) < 0.0) solid[raw_index(c, i + NG3, j + NG3, k + NG3)] = 1;
            }
    std::vector<long> gidx;
    std::vector<double> gimg;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int ir = i + NG3, jr = j + NG3, kr = k + NG3;
                const long q = raw_index(c, ir, jr, kr);
                if (!solid[q]) continue;
                const long nb[6] = {raw_index(c, ir + 1, jr, kr), raw_index(c, ir - 1, jr, kr), raw_index(c, ir, jr + 1, kr),
                                    raw_index(c, ir, jr - 1, kr), raw_index(c, ir, jr, kr + 1), raw_index(c, ir, jr, kr - 1)};
                bool anyfluid = false;
                for (int t = 0; t < 6; t++) anyfluid = anyfluid || !solid[nb[t]];
                if (!anyfluid) continue;   // fully buried
                const double cc[3] = {centroid3(c, 0, ir), centroid3(c, 1, jr), centroid3(c, 2, kr)};
                double best_abs = 1.79769313486231570815e308, best_signed = 0.0, bn[3] = {1.0, 0.0, 0.0};
                for (int b = 0; b < c->n_body; b++) {
                    const double s = body_sd(c, b, cc);
                    if (std::fabs(s) < best_abs) { best_abs = std::fabs(s); best_signed = s; body_normal(c, b, cc, bn); }
                }
                const double push = 2.0 * (-best_signed);
                gidx.push_back(q);
                for (int d = 0; d < 3; d++) gimg.push_back(cc[d] + push * bn[d]);
            }
    std::vector<double> mk;
    const double PI = 3.14159265358979323846;
    if (c->body_kind == 2) {
        // spherical Fibonacci lattice, equal areas 4 pi R^2 / N, radial normals (sphere.rs:213-237)
        const int n = std::max(c->sphere_markers, 12);
        const double golden = PI * (3.0 - std::sqrt(5.0));
        for (int b = 0; b < c->n_body; b++) {
            const double* B = c->body[b];
            const double area = 4.0 * PI * B[3] * B[3] / (double)n;
            for (int m = 0; m < n; m++) {
                const double zr = 1.0 - 2.0 * ((double)m + 0.5) / (double)n;
                const double rr = std::sqrt(std::max(1.0 - zr * zr, 0.0));
                const double phi = golden * (double)m;
                const double dir[3] = {rr * std::cos(phi), rr * std::sin(phi), zr};
                const double row[7] = {B[0] + B[3] * dir[0], B[1] + B[3] * dir[1], B[2] + B[3] * dir[2], dir[0], dir[1], dir[2], area};
                mk.insert(mk.end(), row, row + 7);
            }
        }
    } else {
        // 64 x nz cylinder markers (cfd_ibm/src/lib.rs:566-598)
        const double z_lo = centroid3(c, 2, NG3) - 0.5 * c->h[2];
        const double z_hi = centroid3(c, 2, NG3 + nz - 1) + 0.5 * c->h[2];
        const double lz = z_hi - z_lo;
        const int n_theta = 64, n_z = nz > 1 ? nz : 1;
        for (int b = 0; b < c->n_body; b++) {
            const double* B = c->body[b];
            const double area = 2.0 * PI * B[3] * lz / (double)(n_theta * n_z);
            for (int it = 0; it < n_theta; it++) {
                const double th = 2.0 * PI * ((double)it + 0.5) / (double)n_theta;
                const double nxv = std::cos(th), nyv = std::sin(th);
                for (int iz = 0; iz < n_z; iz++) {
                    const double z = z_lo + ((double)iz + 0.5) * lz / (double)n_z;
                    const double row[7] = {B[0] + B[3] * nxv, B[1] + B[3] * nyv, z, nxv, nyv, 0.0, area};
                    mk.insert(mk.end(), row, row + 7);
                }
            }
        }
    }
    const C5<double> f = {c->freestream[0], c->freestream[1], c->freestream[2], c->freestream[3], c->freestream[4]};
    const double stag[5] = {f.r, 0.0, 0.0, 0.0, f.re - ref::kinetic(f)};
    efvm_status st = EFVM_OK;
    if ((st = alloc3(c, (void**)&c->solid, c->nraw)) != EFVM_OK) return st;
    if ((st = alloc3(c, (void**)&c->ghost_idx, sizeof(long) * std::max<size_t>(1, gidx.size()))) != EFVM_OK) return st;
    if ((st = alloc3(c, (void**)&c->ghost_img, sizeof(double) * std::max<size_t>(3, gimg.size()))) != EFVM_OK) return st;
    if ((st = alloc3(c, (void**)&c->markers, sizeof(double) * std::max<size_t>(7, mk.size()))) != EFVM_OK) return st;
    if ((st = alloc3(c, (void**)&c->mforce, sizeof(double) * 6 * std::max<size_t>(1, mk.size() / 7))) != EFVM_OK) return st;
    if ((st = alloc3(c, (void**)&c->stagnant, sizeof(double) * 5)) != EFVM_OK) return st;
    HIPCHK3(c, hipMemcpy(c->solid, solid.data(), c->nraw, hipMemcpyHostToDevice));
    if (!gidx.empty()) {
        HIPCHK3(c, hipMemcpy(c->ghost_idx, gidx.data(), sizeof(long) * gidx.size(), hipMemcpyHostToDevice));
        HIPCHK3(c, hipMemcpy(c->ghost_img, gimg.data(), sizeof(double) * gimg.size(), hipMemcpyHostToDevice));
    }
    if (!mk.empty()) HIPCHK3(c, hipMemcpy(c->markers, mk.data(), sizeof(double) * mk.size(), hipMemcpyHostToDevice));
    HIPCHK3(c, hipMemcpy(c->stagnant, stag, sizeof(stag), hipMemcpyHostToDevice));
    c->n_ghost = (int)gidx.size();
    c->n_mark = (int)(mk.size() / 7);
    c->ibm_flipped = 0;
    return EFVM_OK;
}

}  // namespace

extern "C" {

efvm_status efvm3_create(const efvm_grid3* grid, const efvm_gas* gas, const efvm_scheme* scheme, efvm_ctx3** out) {
    if (!grid || !gas || !scheme || !out) return EFVM_ERR_ARG;
    *out = nullptr;
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return EFVM_ERR_CONFIG;
    for (int d = 0; d < 3; d++) if (!(grid->hi[d] > grid->lo[d])) return EFVM_ERR_CONFIG;
    if ((long)(grid->nx + 4) * (grid->ny + 4) * (long)(grid->nz + 4) > (1L << 31)) return EFVM_ERR_UNSUPPORTED;
    if (scheme->flux < 0 || scheme->flux > 2 || scheme->recon < 0 || scheme->recon > 1 || scheme->limiter < 0 ||
        scheme->limiter > 1 || scheme->integrator < 0 || scheme->integrator > 1 || scheme->precision < 0 ||
        scheme->precision > 1 || scheme->seam_mode < 0 || scheme->seam_mode > 1)
        return EFVM_ERR_CONFIG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EFVM_ERR_NO_DEVICE;
    efvm_ctx3* c = new (std::nothrow) efvm_ctx3();
    if (!c) return EFVM_ERR_ARG;
    c->grid = *grid; c->gas = *gas; c->sch = *scheme;
    for (int s = 0; s < 6; s++) { c->bc[s].kind = EFVM_BC_OUTFLOW; for (int q = 0; q < 5; q++) c->bc[s].par[q] = 0.0; }
    const int n[3] = {grid->nx, grid->ny, grid->nz};
    for (int d = 0; d < 3; d++) c->h[d] = (grid->hi[d] - grid->lo[d]) / (double)n[d];
    c->NX = grid->nx + 2 * NG3; c->NY = grid->ny + 2 * NG3; c->NZ = grid->nz + 2 * NG3;
    c->nraw = (long)c->NX * c->NY * c->NZ;
    c->ncell = (long)grid->nx * grid->ny * grid->nz;
    c->esz = scheme->precision == EFVM_FP64 ? 8 : 4;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        efvm3_destroy(c);
        return EFVM_ERR_HIP;
    }
    efvm_status st = EFVM_OK;
    for (int s = 0; s < 3 && st == EFVM_OK; s++)
        for (int m = 0; m < 5 && st == EFVM_OK; m++) st = alloc3(c, &c->P[s][m], c->esz * c->nraw);
    if (st == EFVM_OK) st = alloc3(c, (void**)&c->clock, sizeof(DevClock));
    if (st == EFVM_OK) st = alloc3(c, (void**)&c->part, sizeof(double) * 5 * 1024);
    if (st == EFVM_OK) st = alloc3(c, (void**)&c->out, sizeof(double) * 16);
    if (st == EFVM_OK) st = alloc3(c, (void**)&c->yw, sizeof(double) * 3 * grid->ny);
    if (st == EFVM_OK) st = alloc3(c, (void**)&c->zw, sizeof(double) * 3 * grid->nz);
    if (st != EFVM_OK) { efvm3_destroy(c); return st; }
    for (int s = 0; s < 3; s++)
        for (int m = 0; m < 5; m++) (void)hipMemsetAsync(c->P[s][m], 0, c->esz * c->nraw, c->stream);
    if (upload_weights3(c) != EFVM_OK || launch3_clock_set(c->clock, 0.0, 0.0, 0, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        efvm3_destroy(c);
        return EFVM_ERR_HIP;
    }
    *out = c;
    return EFVM_OK;
}

void efvm3_destroy(efvm_ctx3* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph3(c);
    for (int s = 0; s < 3; s++) for (int m = 0; m < 5; m++) if (c->P[s][m]) (void)hipFree(c->P[s][m]);
    if (c->clock) (void)hipFree(c->clock);
    if (c->part) (void)hipFree(c->part);
    if (c->out) (void)hipFree(c->out);
    if (c->yw) (void)hipFree(c->yw);
    if (c->zw) (void)hipFree(c->zw);
    for (int d = 0; d < 3; d++) {
        if (c->width_d[d]) (void)hipFree(c->width_d[d]);
        if (c->cen_d[d]) (void)hipFree(c->cen_d[d]);
    }
    for (int s = 0; s < 6; s++) if (c->bc_tab[s]) (void)hipFree(c->bc_tab[s]);
    free_ibm3(c);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* efvm3_last_error(const efvm_ctx3* c) { return c ? c->err.c_str() : "null context"; }

efvm_status efvm3_set_boundaries(efvm_ctx3* c, const efvm_bc sides[6]) {
    if (!c || !sides) return EFVM_ERR_ARG;
    for (int s = 0; s < 6; s++) {
        if (sides[s].kind < 0 || sides[s].kind > EFVM_BC_TABLE) return fail3(c, EFVM_ERR_CONFIG, "unknown boundary kind");
    }
    for (int ax = 0; ax < 3; ax++)
        if ((sides[2 * ax].kind == EFVM_BC_PERIODIC) != (sides[2 * ax + 1].kind == EFVM_BC_PERIODIC))
            return fail3(c, EFVM_ERR_CONFIG, std::string("periodic boundary on axis ") + std::to_string(ax) +
                                                 " requires both sides periodic (crates/cfd_boundary/src/lib.rs:437-448)");
    for (int s = 0; s < 6; s++) c->bc[s] = sides[s];
    drop_graph3(c);
    // a stretched axis's ghost centroids, and so its viscous weights, follow its periodicity
    return upload_weights3(c);
}

efvm_status efvm3_set_edges(efvm_ctx3* c, int axis, const double* edges, size_t n_edges) {
    if (!c || !edges) return EFVM_ERR_ARG;
    if (axis != 1 && axis != 2)
        return fail3(c, EFVM_ERR_UNSUPPORTED, "only y and z may be stretched: the viscous x term takes one spacing "
                                              "for every cell (crates/cfd_solver/src/lib.rs:479-480)");
    const int n = axis == 1 ? c->grid.ny : c->grid.nz;
    if (n_edges != (size_t)n + 1) return fail3(c, EFVM_ERR_ARG, "edges: n + 1 values along the axis");
    if (edges[0] != c->grid.lo[axis] || edges[n] != c->grid.hi[axis])
        return fail3(c, EFVM_ERR_ARG, "edges: the first and last edge are the grid's lo and hi");
    for (int i = 0; i < n; i++)
        if (!(edges[i + 1] > edges[i])) return fail3(c, EFVM_ERR_ARG, "edges: strictly increasing");
    if (c->n_body > 0) return fail3(c, EFVM_ERR_UNSUPPORTED, "immersed bodies need uniform spacing");
    if (c->rk_stage != 0) return fail3(c, EFVM_ERR_ARG, "the mesh changes at a step boundary only");
    std::vector<double> w(n), cen(n);
    for (int i = 0; i < n; i++) { w[i] = edges[i + 1] - edges[i]; cen[i] = 0.5 * (edges[i] + edges[i + 1]); }
    for (double** p : {&c->width_d[axis], &c->cen_d[axis]})
        if (!*p) {
            efvm_status st = alloc3(c, (void**)p, sizeof(double) * n);
            if (st != EFVM_OK) return st;
        }
    c->edges[axis].assign(edges, edges + n + 1);
    drop_graph3(c);
    HIPCHK3(c, hipMemcpyAsync(c->width_d[axis], w.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK3(c, hipMemcpyAsync(c->cen_d[axis], cen.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    return upload_weights3(c);
}

efvm_status efvm3_set_ghost_table(efvm_ctx3* c, int side, const double* cons5, const unsigned char* given, size_t n) {
    if (!c || !cons5 || !given || side < 0 || side > 5) return EFVM_ERR_ARG;
    if (c->bc[side].kind != EFVM_BC_TABLE) return fail3(c, EFVM_ERR_CONFIG, "ghost table on a side whose kind is not EFVM_BC_TABLE");
    const int axis = side / 2;
    const size_t nf = axis == 0 ? (size_t)c->grid.ny * c->grid.nz
                                : (axis == 1 ? (size_t)c->grid.nx * c->grid.nz : (size_t)c->grid.nx * c->grid.ny);
    if (n != nf)
        return fail3(c, EFVM_ERR_ARG, "ghost table: one entry per boundary face (ny nz for x sides, nx nz for y sides, nx ny for z sides)");
    std::vector<double> h(6 * nf);
    for (size_t k = 0; k < nf; k++) {
        for (int m = 0; m < 5; m++) h[6 * k + m] = cons5[5 * k + m];
        h[6 * k + 5] = given[k] ? 1.0 : 0.0;
    }
    if (!c->bc_tab[side]) {
        efvm_status st = alloc3(c, (void**)&c->bc_tab[side], sizeof(double) * 6 * nf);
        if (st != EFVM_OK) return st;
        drop_graph3(c);   // a captured graph holds arguments without this table
    }
    HIPCHK3(c, hipMemcpyAsync(c->bc_tab[side], h.data(), sizeof(double) * 6 * nf, hipMemcpyHostToDevice, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm3_set_time_step(efvm_ctx3* c, double cfl, double fixed_dt) {
    if (!c || !(cfl > 0.0)) return EFVM_ERR_ARG;
    if (c->rk_stage != 0) return fail3(c, EFVM_ERR_ARG, "time-step control changes at a step boundary only");
    c->sch.cfl = cfl;
    c->sch.fixed_dt = fixed_dt;
    drop_graph3(c);
    return EFVM_OK;
}

efvm_status efvm3_set_body_force(efvm_ctx3* c, double gx) {
    if (!c) return EFVM_ERR_ARG;
    c->body_force = 1; c->gx = gx;
    drop_graph3(c);
    return EFVM_OK;
}

efvm_status efvm3_set_sponge(efvm_ctx3* c, double x0, double x1, double sigma, const double target[5]) {
    if (!c || !target || !(x1 > x0)) return EFVM_ERR_ARG;
    if (c->sch.integrator != EFVM_INTEGRATOR_RK3) return fail3(c, EFVM_ERR_UNSUPPORTED, "sponge needs RK3");
    c->sponge = 1;
    std::memset(&c->sp, 0, sizeof(c->sp));
    c->sp.x0 = x0; c->sp.x1 = x1; c->sp.sigma = sigma; c->sp.uniform = 1;
    for (int m = 0; m < 5; m++) c->sp.target[m] = target[m];
    drop_graph3(c);
    return EFVM_OK;
}

efvm_status efvm3_upload_state(efvm_ctx3* c, const double* cons5, size_t n) {
    if (!c || !cons5) return EFVM_ERR_ARG;
    const bool raw = (long)n == c->nraw;
    if (!raw && (long)n != c->ncell) return fail3(c, EFVM_ERR_ARG, "upload: n_cells must be nx*ny*nz or the raw padded count");
    const int nx = c->grid.nx, ny = c->grid.ny, nz = c->grid.nz;
    std::vector<char> h[5];
    for (int m = 0; m < 5; m++) h[m].resize(c->esz * c->nraw);
    for (int k = 0; k < c->NZ; k++)
        for (int j = 0; j < c->NY; j++)
            for (int i = 0; i < c->NX; i++) {
                const long q = raw_index(c, i, j, k);
                long src;
                if (raw) {
                    src = q;
                } else {   // ghost cells take their nearest interior cell
                    const int ii = std::min(std::max(i - NG3, 0), nx - 1), jj = std::min(std::max(j - NG3, 0), ny - 1),
                              kk = std::min(std::max(k - NG3, 0), nz - 1);
                    src = ii + (long)nx * (jj + (long)ny * kk);
                }
                for (int m = 0; m < 5; m++) {
                    const double v = cons5[5 * src + m];
                    if (c->esz == 8) reinterpret_cast<double*>(h[m].data())[q] = v;
                    else reinterpret_cast<float*>(h[m].data())[q] = (float)v;
                }
            }
    c->rk_stage = 0;
    for (int m = 0; m < 5; m++)
        HIPCHK3(c, hipMemcpyAsync(c->P[c->cur][m], h[m].data(), c->esz * c->nraw, hipMemcpyHostToDevice, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    c->have_prev = 0;
    c->rhs_state = 0;
    return EFVM_OK;
}

efvm_status efvm3_download_state(efvm_ctx3* c, double* cons5, size_t n) {
    if (!c || !cons5) return EFVM_ERR_ARG;
    const bool raw = (long)n == c->nraw;
    if (!raw && (long)n != c->ncell) return fail3(c, EFVM_ERR_ARG, "download: n_cells must be nx*ny*nz or the raw padded count");
    const int set = current_set(c);
    std::vector<char> h[5];
    for (int m = 0; m < 5; m++) {
        h[m].resize(c->esz * c->nraw);
        HIPCHK3(c, hipMemcpyAsync(h[m].data(), c->P[set][m], c->esz * c->nraw, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    const int nx = c->grid.nx, ny = c->grid.ny, nz = c->grid.nz;
    auto val = [&](int m, long q) {
        return c->esz == 8 ? reinterpret_cast<double*>(h[m].data())[q] : (double)reinterpret_cast<float*>(h[m].data())[q];
    };
    if (raw) {
        for (long q = 0; q < c->nraw; q++) for (int m = 0; m < 5; m++) cons5[5 * q + m] = val(m, q);
    } else {
        long o = 0;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++, o++) {
                    const long q = raw_index(c, i + NG3, j + NG3, k + NG3);
                    for (int m = 0; m < 5; m++) cons5[5 * o + m] = val(m, q);
                }
    }
    return EFVM_OK;
}

efvm_status efvm3_get_clock(efvm_ctx3* c, double* dt, double* time, long* step, int* stage) {
    if (!c) return EFVM_ERR_ARG;
    DevClock h;
    HIPCHK3(c, hipMemcpyAsync(&h, c->clock, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    if (dt) *dt = h.dt;
    if (time) *time = h.time;
    if (step) *step = (long)h.step;
    if (stage) *stage = c->rk_stage;
    if (h.bad) return fail3(c, EFVM_ERR_NONPHYSICAL, "CFL scan saw " + std::to_string(h.bad) + " non-finite or non-positive-density cells");
    return EFVM_OK;
}

efvm_status efvm3_set_clock(efvm_ctx3* c, double time, long step) {
    if (!c) return EFVM_ERR_ARG;
    HIPCHK3(c, launch3_clock_set(c->clock, 0.0, time, step, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    c->rk_stage = 0; c->parity = 0;
    return EFVM_OK;
}

efvm_status efvm3_cfl(efvm_ctx3* c, double* dt_out) {
    if (!c || !dt_out) return EFVM_ERR_ARG;
    if (c->sch.fixed_dt > 0.0) { *dt_out = c->sch.fixed_dt; return EFVM_OK; }
    Args3 a = args3(c);
    const int set = current_set(c);
    planes3(a, c->P[set], c->P[set], c->P[set]);
    LAUNCH3(c, launch3_cfl(a, (int)c->esz, c->part, &c->clock->lam_tmp, c->stream), "cfl");
    double lam = 0.0;
    HIPCHK3(c, hipMemcpyAsync(&lam, &c->clock->lam_tmp, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    *dt_out = lam > 0.0 ? c->sch.cfl / lam : INFINITY;
    return EFVM_OK;
}

efvm_status efvm3_run(efvm_ctx3* c) {
    if (!c) return EFVM_ERR_ARG;
    return run3(c);
}

efvm_status efvm3_step(efvm_ctx3* c, int n_steps) {
    if (!c || n_steps < 0) return EFVM_ERR_ARG;
    HIPCHK3(c, hipEventRecord(c->ev0, c->stream));
    const int nsub = c->sch.integrator == EFVM_INTEGRATOR_RK3 ? 3 : 1;
    for (int k = 0; k < n_steps; k++)
        for (int q = 0; q < nsub; q++) {
            efvm_status st = run3(c);
            if (st != EFVM_OK) return st;
        }
    HIPCHK3(c, hipEventRecord(c->ev1, c->stream));
    c->timed = 1;
    return EFVM_OK;
}

efvm_status efvm3_synchronize(efvm_ctx3* c) {
    if (!c) return EFVM_ERR_ARG;
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm3_graph_capture(efvm_ctx3* c, int n_steps) {
    if (!c || n_steps <= 0 || (n_steps & 1)) return fail3(c, EFVM_ERR_ARG, "graph capture needs an even positive step count");
    if (c->rk_stage != 0) return fail3(c, EFVM_ERR_ARG, "graph capture must start at a step boundary");
    // the first step after a body is added scans the CFL before the mask refresh flips the cells
    // inside it; a replayed graph cannot repeat that order, so that step runs eagerly
    if (c->n_body > 0 && !c->ibm_flipped)
        return fail3(c, EFVM_ERR_ARG, "graph capture after a body is added needs one eager step first");
    drop_graph3(c);
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    const int cur0 = c->cur, prev0 = c->prev, par0 = c->parity, have0 = c->have_prev, rhs0 = c->rhs_state;
    HIPCHK3(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    efvm_status st = EFVM_OK;
    const int nsub = c->sch.integrator == EFVM_INTEGRATOR_RK3 ? 3 : 1;
    for (int k = 0; k < n_steps && st == EFVM_OK; k++)
        for (int q = 0; q < nsub && st == EFVM_OK; q++) st = run3(c);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st != EFVM_OK) { if (g) (void)hipGraphDestroy(g); return st; }
    if (e != hipSuccess) return fail3(c, EFVM_ERR_HIP, std::string("end capture: ") + hipGetErrorString(e));
    c->graph = g;
    HIPCHK3(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
    c->cur = cur0; c->prev = prev0; c->parity = par0; c->rk_stage = 0; c->have_prev = have0; c->rhs_state = rhs0;
    c->graph_cur = cur0; c->graph_parity = par0;
    return EFVM_OK;
}

efvm_status efvm3_graph_launch(efvm_ctx3* c) {
    if (!c || !c->graph_exec) return fail3(c, EFVM_ERR_ARG, "no captured graph");
    // the captured steps start from the plane set and CFL slot of the capture (each step scans
    // its own CFL, so an upload needs nothing more): refuse to start mid-step or elsewhere
    if (c->rk_stage != 0) return fail3(c, EFVM_ERR_ARG, "graph replay must start at a step boundary");
    if (c->cur != c->graph_cur || c->parity != c->graph_parity)
        return fail3(c, EFVM_ERR_ARG, "the state is on another plane set or CFL slot than at capture (an odd number "
                                      "of eager steps since): capture the graph again");
    c->have_prev = 1;
    c->rhs_state = 1;
    HIPCHK3(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK3(c, hipGraphLaunch(c->graph_exec, c->stream));
    HIPCHK3(c, hipEventRecord(c->ev1, c->stream));
    c->timed = 1;
    return EFVM_OK;
}

efvm_status efvm3_last_elapsed_ms(efvm_ctx3* c, double* ms) {
    if (!c || !ms || !c->timed) return EFVM_ERR_ARG;
    HIPCHK3(c, hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK3(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return EFVM_OK;
}

efvm_status efvm3_ibm_add_sphere(efvm_ctx3* c, double cx, double cy, double cz, double r, const double fs[5], int n_markers) {
    if (!c || !fs || !(r > 0.0) || n_markers < 1) return EFVM_ERR_ARG;
    if (c->body_kind == 1) return fail3(c, EFVM_ERR_UNSUPPORTED, "a context holds spheres or cylinders, not both");
    if (!c->edges[1].empty() || !c->edges[2].empty()) return fail3(c, EFVM_ERR_UNSUPPORTED, "immersed bodies need uniform spacing");
    if (c->n_body >= 8) return fail3(c, EFVM_ERR_UNSUPPORTED, "at most 8 bodies");
    if (c->rk_stage != 0) return fail3(c, EFVM_ERR_ARG, "bodies are added at a step boundary");
    c->body_kind = 2;
    c->body[c->n_body][0] = cx; c->body[c->n_body][1] = cy; c->body[c->n_body][2] = cz; c->body[c->n_body][3] = r;
    c->n_body++;
    c->sphere_markers = n_markers;
    std::memcpy(c->freestream, fs, sizeof(double) * 5);
    return build_ibm3(c);
}

efvm_status efvm3_ibm_add_cylinder(efvm_ctx3* c, double cx, double cy, double r, const double fs[5]) {
    if (!c || !fs || !(r > 0.0)) return EFVM_ERR_ARG;
    if (c->body_kind == 2) return fail3(c, EFVM_ERR_UNSUPPORTED, "a context holds spheres or cylinders, not both");
    if (!c->edges[1].empty() || !c->edges[2].empty()) return fail3(c, EFVM_ERR_UNSUPPORTED, "immersed bodies need uniform spacing");
    if (c->n_body >= 8) return fail3(c, EFVM_ERR_UNSUPPORTED, "at most 8 bodies");
    if (c->rk_stage != 0) return fail3(c, EFVM_ERR_ARG, "bodies are added at a step boundary");
    c->body_kind = 1;
    c->body[c->n_body][0] = cx; c->body[c->n_body][1] = cy; c->body[c->n_body][2] = 0.0; c->body[c->n_body][3] = r;
    c->n_body++;
    std::memcpy(c->freestream, fs, sizeof(double) * 5);
    return build_ibm3(c);
}

efvm_status efvm3_marker_force(efvm_ctx3* c, double out[6]) {
    if (!c || !out) return EFVM_ERR_ARG;
    if (c->n_body == 0 || c->n_mark == 0) { for (int d = 0; d < 6; d++) out[d] = 0.0; return EFVM_OK; }
    Args3 a = args3(c);
    const int set = current_set(c);
    planes3(a, c->P[set], c->P[set], c->P[set]);
    LAUNCH3(c, launch3_markers(a, (int)c->esz, c->n_mark, c->markers, c->h[0], c->mforce, c->out, c->stream), "marker force");
    HIPCHK3(c, hipMemcpyAsync(out, c->out, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

efvm_status efvm3_wall_force(efvm_ctx3* c, double out[6]) {
    if (!c || !out) return EFVM_ERR_ARG;
    // the mask takes the bodies on the first substage (the reference's Setup phase): no solid cell before
    if (c->n_body == 0 || !c->ibm_flipped) { for (int d = 0; d < 6; d++) out[d] = 0.0; return EFVM_OK; }
    Args3 a = args3(c);
    const int set = current_set(c);
    planes3(a, c->P[set], c->P[set], c->P[set]);
    LAUNCH3(c, launch3_wall(a, (int)c->esz, c->n_ghost, c->ghost_idx, c->out, c->stream), "wall force");
    HIPCHK3(c, hipMemcpyAsync(out, c->out, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

// CFDRST01 (examples/restart_determinism/main.rs:177-226): "CFDRST01", dt, time (f64), step
// (u64), rk stage (u8), then the columns u, rhs, u0, each a u64 cell count and 5 f64 per cell
// over the raw padded array, little-endian.  The 3-D context stores that array, ghost cells
// included, so every column is the reference's:
//   u    the current substage's state;
//   rhs  the last substage's right-hand side (crates/cfd_solver/src/lib.rs:79-88 clears it only
//        at the next PreFlux), recomputed from that substage's input set, whose ghost and IBM
//        mirror cells the substage filled in place; zero before any substage;
//   u0   save_state's copy taken at the first substage (crates/cfd_solver/src/integrate.rs:76-78):
//        U^n mid-step, the state before the completed step at a step boundary (kept by the
//        plane-set rotation); zero before the first step and for Euler, which never saves.
static efvm_status read_set3(efvm_ctx3* c, int set, double* dst) {
    const long nraw = c->nraw;
    std::vector<char> h[5];
    for (int m = 0; m < 5; m++) {
        h[m].resize(c->esz * nraw);
        HIPCHK3(c, hipMemcpyAsync(h[m].data(), c->P[set][m], c->esz * nraw, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    for (long q = 0; q < nraw; q++)
        for (int m = 0; m < 5; m++)
            dst[5 * q + m] = c->esz == 8 ? reinterpret_cast<double*>(h[m].data())[q]
                                         : (double)reinterpret_cast<float*>(h[m].data())[q];
    return EFVM_OK;
}

static efvm_status rhs_column3(efvm_ctx3* c, double* dst) {
    const long nraw = c->nraw;
    if (c->rhs_state == 0) { std::fill(dst, dst + 5 * nraw, 0.0); return EFVM_OK; }
    if (c->rhs_state == 2) { std::copy(c->loaded_rhs.begin(), c->loaded_rhs.end(), dst); return EFVM_OK; }
    const bool rk3 = c->sch.integrator == EFVM_INTEGRATOR_RK3;
    // input of the last substage: U2 (the third set) after stage 3, U^{n} (now prev) after an
    // Euler step, U^n (cur) after stage 1, U1 (prev) after stage 2
    int src = c->prev;
    if (c->rk_stage == 0 && rk3) src = 3 - c->cur - c->prev;
    if (c->rk_stage == 1) src = c->cur;
    void* tmp[5] = {};
    efvm_status st = EFVM_OK;
    for (int m = 0; m < 5 && st == EFVM_OK; m++) st = alloc3(c, &tmp[m], c->esz * nraw);
    if (st == EFVM_OK) {
        Args3 a = args3(c);
        a.parity = c->parity;
        a.stage = 0;
        a.rhs_only = 1;
        planes3(a, c->P[src], c->P[src], tmp);
        hipError_t e = launch3_stage(a, (int)c->esz, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = fail3(c, EFVM_ERR_HIP, std::string("rhs column: ") + hipGetErrorString(e));
    }
    if (st == EFVM_OK) {
        std::vector<char> h[5];
        for (int m = 0; m < 5 && st == EFVM_OK; m++) {
            h[m].resize(c->esz * nraw);
            if (hipMemcpy(h[m].data(), tmp[m], c->esz * nraw, hipMemcpyDeviceToHost) != hipSuccess)
                st = fail3(c, EFVM_ERR_HIP, "rhs column copy");
        }
        if (st == EFVM_OK)
            for (long q = 0; q < nraw; q++)
                for (int m = 0; m < 5; m++)
                    dst[5 * q + m] = c->esz == 8 ? reinterpret_cast<double*>(h[m].data())[q]
                                                 : (double)reinterpret_cast<float*>(h[m].data())[q];
    }
    for (int m = 0; m < 5; m++) if (tmp[m]) (void)hipFree(tmp[m]);
    return st;
}

efvm_status efvm3_checkpoint_save(efvm_ctx3* c, const char* path) {
    if (!c || !path) return EFVM_ERR_ARG;
    const long nraw = c->nraw;
    std::vector<double> u(5 * nraw), rhs(5 * nraw), u0(5 * nraw, 0.0);
    efvm_status st = efvm3_download_state(c, u.data(), nraw);    // the current substage's state
    if (st != EFVM_OK) return st;
    if ((st = rhs_column3(c, rhs.data())) != EFVM_OK) return st;
    const bool rk3 = c->sch.integrator == EFVM_INTEGRATOR_RK3;
    const int u0_set = !rk3 ? -1 : (c->rk_stage != 0 ? c->cur : (c->have_prev ? c->prev : -1));
    if (u0_set >= 0 && (st = read_set3(c, u0_set, u0.data())) != EFVM_OK) return st;
    double dt = 0.0, time = 0.0;
    long step = 0;
    int stage = 0;
    if ((st = efvm3_get_clock(c, &dt, &time, &step, &stage)) != EFVM_OK) return st;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail3(c, EFVM_ERR_ARG, std::string("checkpoint: cannot open ") + path);
    const unsigned long long step_u = (unsigned long long)step, len = (unsigned long long)nraw;
    const unsigned char stage_u = (unsigned char)stage;
    bool ok = std::fwrite("CFDRST01", 1, 8, f) == 8 && std::fwrite(&dt, 8, 1, f) == 1 && std::fwrite(&time, 8, 1, f) == 1 &&
              std::fwrite(&step_u, 8, 1, f) == 1 && std::fwrite(&stage_u, 1, 1, f) == 1;
    const double* cols[3] = {u.data(), rhs.data(), u0.data()};
    for (int k = 0; k < 3 && ok; k++)
        ok = std::fwrite(&len, 8, 1, f) == 1 && std::fwrite(cols[k], sizeof(double), 5 * nraw, f) == (size_t)(5 * nraw);
    const int cl = std::fclose(f);
    if (!ok || cl != 0) return fail3(c, EFVM_ERR_ARG, std::string("checkpoint: short write to ") + path);
    return EFVM_OK;
}

// Load a step-boundary CFDRST01 file of the same grid: the u column (ghosts included) becomes
// the state, the clock continues from its dt, time and step, and the next step recomputes the
// CFL, so the run continues bit for bit (crates/cfd_solver/tests/restart.rs:86-105).  Bodies
// keep the context's mask state: on a context that has not run, the first substage refreshes
// the mask as a new App does.
efvm_status efvm3_checkpoint_load(efvm_ctx3* c, const char* path) {
    if (!c || !path) return EFVM_ERR_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail3(c, EFVM_ERR_ARG, std::string("checkpoint: cannot open ") + path);
    char magic[8];
    double dt = 0.0, time = 0.0;
    unsigned long long step = 0, len = 0;
    unsigned char stage = 0;
    const long nraw = c->nraw;
    std::vector<double> u(5 * nraw), rhs(5 * nraw), u0(5 * nraw);
    efvm_status st = EFVM_OK;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "CFDRST01", 8) != 0) st = EFVM_ERR_CONFIG;
    else if (std::fread(&dt, 8, 1, f) != 1 || std::fread(&time, 8, 1, f) != 1 || std::fread(&step, 8, 1, f) != 1 ||
             std::fread(&stage, 1, 1, f) != 1)
        st = EFVM_ERR_CONFIG;
    else if (stage > 2) st = EFVM_ERR_CONFIG;
    else if (stage != 0) st = EFVM_ERR_UNSUPPORTED;
    for (int k = 0; k < 3 && st == EFVM_OK; k++) {
        if (std::fread(&len, 8, 1, f) != 1 || (long)len != nraw) { st = EFVM_ERR_CONFIG; break; }
        double* dst = k == 0 ? u.data() : (k == 1 ? rhs.data() : u0.data());
        if (std::fread(dst, sizeof(double), 5 * nraw, f) != (size_t)(5 * nraw)) st = EFVM_ERR_CONFIG;
    }
    char extra;
    if (st == EFVM_OK && std::fread(&extra, 1, 1, f) != 0) st = EFVM_ERR_CONFIG;
    std::fclose(f);
    if (st == EFVM_ERR_UNSUPPORTED) return fail3(c, st, "checkpoint: restarts are taken at a step boundary (rk stage 0)");
    if (st != EFVM_OK) return fail3(c, st, "checkpoint: not a CFDRST01 file of this grid (magic, header, cell counts or length)");
    if ((st = efvm3_upload_state(c, u.data(), nraw)) != EFVM_OK) return st;
    // u0 (the state before the file's last step) goes to the prev set and rhs is kept on the host:
    // a checkpoint written before the next substage carries both unchanged
    {
        std::vector<char> h[5];
        for (int m = 0; m < 5; m++) {
            h[m].resize(c->esz * nraw);
            for (long q = 0; q < nraw; q++) {
                const double v = u0[5 * q + m];
                if (c->esz == 8) reinterpret_cast<double*>(h[m].data())[q] = v;
                else reinterpret_cast<float*>(h[m].data())[q] = (float)v;
            }
            HIPCHK3(c, hipMemcpyAsync(c->P[c->prev][m], h[m].data(), c->esz * nraw, hipMemcpyHostToDevice, c->stream));
        }
        HIPCHK3(c, hipStreamSynchronize(c->stream));
    }
    c->have_prev = 1;
    c->loaded_rhs.swap(rhs);
    c->rhs_state = 2;
    drop_graph3(c);
    HIPCHK3(c, launch3_clock_set(c->clock, dt, time, (long long)step, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    c->parity = 0;
    return EFVM_OK;
}

efvm_status efvm3_write_vtk(const efvm_grid3* g, const efvm_gas* gas, const double* cons5, size_t n_cells, const char* path) {
    if (!g || !gas || !cons5 || !path || g->nx < 1 || g->ny < 1 || g->nz < 1) return EFVM_ERR_ARG;
    if ((long)n_cells != (long)g->nx * g->ny * g->nz) return EFVM_ERR_ARG;
    const int nn[3] = {g->nx, g->ny, g->nz};
    double lo[3], h[3];
    for (int d = 0; d < 3; d++) { lo[d] = g->lo[d]; h[d] = (g->hi[d] - g->lo[d]) / (double)nn[d]; }
    return vtk_write_cells(nn, lo, h, gas->gamma, gas->r_gas, cons5, path, nullptr);
}

efvm_status efvm3_write_vtk_edges(const efvm_grid3* g, const double* y_edges, const double* z_edges, const efvm_gas* gas,
                                  const double* cons5, size_t n_cells, const char* path) {
    if (!g || !gas || !cons5 || !path || g->nx < 1 || g->ny < 1 || g->nz < 1) return EFVM_ERR_ARG;
    if ((long)n_cells != (long)g->nx * g->ny * g->nz) return EFVM_ERR_ARG;
    const int nn[3] = {g->nx, g->ny, g->nz};
    const double* edges[3] = {nullptr, y_edges, z_edges};
    for (int d = 1; d < 3; d++)
        if (edges[d]) {
            if (edges[d][0] != g->lo[d] || edges[d][nn[d]] != g->hi[d]) return EFVM_ERR_ARG;
            for (int i = 0; i < nn[d]; i++) if (!(edges[d][i + 1] > edges[d][i])) return EFVM_ERR_ARG;
        }
    double lo[3], h[3];
    for (int d = 0; d < 3; d++) { lo[d] = g->lo[d]; h[d] = (g->hi[d] - g->lo[d]) / (double)nn[d]; }
    return vtk_write_cells(nn, lo, h, gas->gamma, gas->r_gas, cons5, path, edges);
}

efvm_status efvm3_conservation_sums(efvm_ctx3* c, double out[5]) {
    if (!c || !out) return EFVM_ERR_ARG;
    Args3 a = args3(c);
    const int set = current_set(c);
    planes3(a, c->P[set], c->P[set], c->P[set]);
    // a stretched mesh weighs each cell by its own volume inside the kernel
    const bool stretched = c->width_d[1] || c->width_d[2];
    const double vol = stretched ? 1.0 : c->h[0] * c->h[1] * c->h[2];
    const int nb = (int)std::min<long>(1024, (c->ncell + 255) / 256);
    LAUNCH3(c, launch3_sums(a, (int)c->esz, c->part, nb, vol, c->out + 8, c->stream), "sums");
    HIPCHK3(c, hipMemcpyAsync(out, c->out + 8, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    HIPCHK3(c, hipStreamSynchronize(c->stream));
    return EFVM_OK;
}

}  // extern "C"
