/*
 * efvm.h -- C ABI of the MI355X explicit compressible finite-volume update.
 *
 * This is the drop-in boundary for the reference's FIELD/GRASS plugin path of
 * SueHeir/dev_field_efvm: a host-side plugin (Rust `GpuSolverPlugin`, or the
 * C++/Python mirrors shipped here) registers the same schedule phases as
 * SolverPlugin + FluxPlugin/MusclPlugin + IntegratorPlugin + BoundaryPlugin +
 * GhostCellIbmPlugin/CylinderIbmPlugin and calls these entry points instead of
 * running the CPU systems.  Plain C types only: pointers, sizes, doubles, ints.
 *
 * State wire order is the reference's `Component` serialisation of `ConsVar`:
 * 5 doubles per cell (rho, rho_u, rho_v, rho_w, rho_E)
 * (crates/cfd_state/src/lib.rs:54-76), interior cells only, x fastest.  The
 * efvm_* context is the 2-D path (the reference's nz = 1 UniformMesh with z slip
 * walls): rho_w must be 0 on upload and is 0 on download.  The efvm3_* context
 * at the end of this header takes any nz and all five components.
 *
 * Errors: every call returns an efvm_status; no C++ exception or abort crosses
 * the ABI (the reference panics, e.g. crates/cfd_solver/src/lib.rs:99,
 * crates/cfd_boundary/src/lib.rs:516-518).  efvm_last_error() gives the text.
 * One context is not thread-safe; separate contexts may be used concurrently.
 */
#ifndef EFVM_H
#define EFVM_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EFVM_OK = 0,
    EFVM_ERR_ARG = 1,          /* bad argument (null pointer, size mismatch)          */
    EFVM_ERR_CONFIG = 2,       /* invalid configuration (half-periodic axis, ...)     */
    EFVM_ERR_HIP = 3,          /* HIP runtime error                                   */
    EFVM_ERR_UNSUPPORTED = 4,  /* feature outside the 2-D GPU path (e.g. rho_w != 0)  */
    EFVM_ERR_NONPHYSICAL = 5,  /* CFL scan saw non-finite or non-positive density     */
    EFVM_ERR_NO_DEVICE = 6     /* no HIP device visible                               */
} efvm_status;

/* Flux schemes: cfd_riemann Rusanov / Hllc (with Thornber) / HllcTextbook
 * (crates/cfd_riemann/src/lib.rs:213-441). */
enum { EFVM_FLUX_RUSANOV = 0, EFVM_FLUX_HLLC = 1, EFVM_FLUX_HLLC_TEXTBOOK = 2 };
/* Reconstruction: FluxPlugin (first order) or MusclPlugin with a SlopeLimiter
 * (crates/cfd_solver/src/lib.rs:159-163,799-842). */
enum { EFVM_RECON_FIRST_ORDER = 0, EFVM_RECON_MUSCL = 1 };
enum { EFVM_LIMITER_MINMOD = 0, EFVM_LIMITER_MC = 1 };
enum { EFVM_INTEGRATOR_EULER = 0, EFVM_INTEGRATOR_RK3 = 1 };   /* integrate.rs:46-105 */
enum { EFVM_FP64 = 0, EFVM_FP32 = 1 };
/* Viscosity model, crates/cfd_eos/src/lib.rs:100-117. */
enum { EFVM_VISC_SUTHERLAND = 0, EFVM_VISC_CONSTANT = 1 };
/* Periodic-seam treatment: 0 = the reference's (ghost side of a seam face is not
 * reconstructed, crates/cfd_solver/src/lib.rs:327-331); 1 = conservative. */
enum { EFVM_SEAM_REFERENCE = 0, EFVM_SEAM_CONSERVATIVE = 1 };
/* Boundary conditions, crates/cfd_boundary/src/lib.rs:173-362.  Sides are
 * XLo=0, XHi=1, YLo=2, YHi=3 (BoundarySide order; z sides are slip walls). */
enum {
    EFVM_BC_OUTFLOW = 0,            /* SupersonicOutflow (registry default)     */
    EFVM_BC_PERIODIC = 1,           /* Periodic (both sides of an axis)         */
    EFVM_BC_SLIP = 2,               /* SlipWall                                 */
    EFVM_BC_NOSLIP = 3,             /* NoSlipWall                               */
    EFVM_BC_SYMMETRY = 4,           /* Symmetry                                 */
    EFVM_BC_SUPERSONIC_INFLOW = 5,  /* par = rho, u, v, w, p                    */
    EFVM_BC_SUBSONIC_INFLOW = 6,    /* par = rho, u, v, w, T                    */
    EFVM_BC_PRESSURE_OUTLET = 7,    /* par[0] = p                               */
    EFVM_BC_TABLE = 8               /* FromOutside: ghost states supplied per face
                                       (efvm_set_ghost_table), else outflow copy */
};

/* UniformMesh {nx, ny, nz=1, ng=2, bounds_lo, bounds_hi}: x in [lo_x, hi_x],
 * y in [lo_y, hi_y], z in [0, dz]. */
typedef struct { int nx, ny; double lo_x, lo_y, hi_x, hi_y, dz; } efvm_grid;
/* IdealGas {gamma, r_gas} + Viscosity. */
typedef struct { double gamma, r_gas; int viscosity; double mu; } efvm_gas;
/* SolverConfig {cfl, fixed_dt, viscous} + flux/reconstruction/integrator choice. */
typedef struct {
    int flux, recon, limiter, integrator, precision, viscous, seam_mode;
    double cfl, fixed_dt;   /* fixed_dt <= 0: CFL-limited */
} efvm_scheme;
typedef struct { int kind; double par[5]; } efvm_bc;

typedef struct efvm_ctx efvm_ctx;

/* Context = one App with its mesh, state columns and plugins
 * (replaces FieldDefaultPlugins + CfdStatePlugin + SolverPlugin + flux plugin). */
efvm_status efvm_ctx_create(const efvm_grid* grid, const efvm_gas* gas, const efvm_scheme* scheme, efvm_ctx** out);
void efvm_ctx_destroy(efvm_ctx* ctx);
const char* efvm_last_error(const efvm_ctx* ctx);
/* Tile shape of the stage kernel (rows per wave strip, waves per workgroup);
 * 0 keeps the current value.  Used by the autotuner. */
efvm_status efvm_set_launch(efvm_ctx* ctx, int strip_rows, int waves_per_block);
/* The current launch shape and the stage kernel the context runs: *kernel = 0 one-column
 * strips (64 columns per wave), 1 two-column strips (128 per wave), 2 FP64 strips, 3 learned
 * face kernels; *pair_fits = which two-column form applies (FP32, no boundary features):
 * 0 none, 1 the periodic form (both axes periodic, nx even), 2 the general form (periodic,
 * outflow, slip, no-slip and symmetry sides; nx even on a periodic x axis).  Any pointer may
 * be null.  No reference counterpart (a tuning query). */
efvm_status efvm_get_launch(const efvm_ctx* ctx, int* strip_rows, int* waves_per_block, int* kernel, int* pair_fits);
/* The gcnArchName of HIP device `device` (e.g. "gfx950:sramecc+:xnack-") into arch[n]: the
 * autotuner's cache key.  No reference counterpart. */
efvm_status efvm_device_arch(int device, char* arch, size_t n);
/* Use a two-column strip kernel where one applies (1, the default) or always the one-column
 * kernel (0).  The periodic form takes grids periodic on both axes with nx even; the general
 * form takes any mix of periodic, outflow, slip, no-slip and symmetry sides (nx even on a
 * periodic x axis) with no boundary features.  Results are bitwise the same.  Used by the
 * autotuner; no reference counterpart. */
efvm_status efvm_set_pair_kernel(efvm_ctx* ctx, int allow);
/* Run RK3 stages 2 and 3 of each whole step as one fused launch where it applies (allow = 1, the
 * default): FP32, the periodic two-column path (both axes periodic, nx even, no boundary
 * features), ny >= 8.  A workgroup of two waves per strip runs stage 2 and stage 3, U2 passing
 * between them through LDS, so each step moves 80 B per cell instead of 128; results are
 * bitwise those of three substage launches.  strip_rows sets the fused strips' height (0 keeps
 * the current value).  Steps started mid-step (after run() substages) and run() itself use the
 * separate stage launches.  Used by the autotuner; no reference counterpart
 * (crates/cfd_solver/src/integrate.rs:69-101 runs the three stages as three passes). */
efvm_status efvm_set_fused_step(efvm_ctx* ctx, int allow, int strip_rows);
/* The fused step's setting and strip height, and whether this context's configuration admits it
 * (*fits).  Any pointer may be null.  No reference counterpart (a tuning query). */
efvm_status efvm_get_fused_step(const efvm_ctx* ctx, int* allow, int* strip_rows, int* fits);

/* BoundaryRegistry + StructuredBoundaryPlugin (crates/cfd_boundary/src/lib.rs:369-533):
 * sides[4] = XLo, XHi, YLo, YHi.  A half-periodic axis is EFVM_ERR_CONFIG. */
efvm_status efvm_set_boundaries(efvm_ctx* ctx, const efvm_bc sides[4]);
/* FromOutside + BoundaryInputs (crates/cfd_boundary/src/lib.rs:108-147; the reference test
 * crates/cfd_solver/tests/boundary_face.rs:96-118): ghost states another code supplies for
 * the faces of a side whose kind is EFVM_BC_TABLE.  n = ny for an x side, nx for a y side;
 * cons5 holds n states in the wire order (rho_w = 0); given[k] = 0 leaves face k to the
 * fallback, a copy of the interior cell (SupersonicOutflow).  A position-dependent FnBoundary
 * is the same table filled from the face centroids.  May be called again between steps
 * (also with a captured graph: the table is read in place). */
efvm_status efvm_set_ghost_table(efvm_ctx* ctx, int side, const double* cons5, const unsigned char* given, size_t n);
/* Case systems of examples/cylinder_drag_validation/main.rs:151-229. */
efvm_status efvm_set_parabolic_inflow(efvm_ctx* ctx, double u_max, double height, double rho);
efvm_status efvm_set_sponge(efvm_ctx* ctx, double x_start, double x_end, double sigma, double u_max, double height,
                            double rho, double p);
/* SolverConfig {cfl, fixed_dt} after creation (crates/cfd_solver/src/lib.rs:39-58):
 * fixed_dt > 0 steps with that dt, else dt = cfl / lambda_max.  At a step boundary. */
efvm_status efvm_set_time_step(efvm_ctx* ctx, double cfl, double fixed_dt);
/* Uniform x body force (crates/cfd_solver/tests/poiseuille_channel_validation.rs:105-120). */
efvm_status efvm_set_body_force(efvm_ctx* ctx, double gx);

/* State transfer, CfdState.u in Component wire order (interior cells, x fastest).
 * Synchronous.  Upload also resets the CFL so the next step recomputes dt. */
efvm_status efvm_upload_state(efvm_ctx* ctx, const double* cons5, size_t n_cells);
efvm_status efvm_download_state(efvm_ctx* ctx, double* cons5, size_t n_cells);

/* SolverState {dt, time, step, rk_stage} (crates/cfd_solver/src/lib.rs:61-73). */
efvm_status efvm_get_clock(efvm_ctx* ctx, double* dt, double* time, long* step, int* rk_stage);
efvm_status efvm_set_clock(efvm_ctx* ctx, double time, long step);
/* cfl_system (crates/cfd_solver/src/lib.rs:626-677): dt of the current state, FP64. */
efvm_status efvm_cfl(efvm_ctx* ctx, double* dt_out);

/* One app.run() = one RK substage through every phase (Setup .. Output). */
efvm_status efvm_run(efvm_ctx* ctx);
/* n full time steps, enqueued without host synchronisation (dt stays on device). */
efvm_status efvm_step(efvm_ctx* ctx, int n_steps);
/* The harness loop `loop { if time >= t_final {break}; step }`. */
efvm_status efvm_advance_to(efvm_ctx* ctx, double t_final, long max_steps, long* steps_done);
efvm_status efvm_synchronize(efvm_ctx* ctx);

/* hipGraph of `n_steps` whole steps (n_steps even: the CFL slots alternate by
 * step parity); launch replays it. */
efvm_status efvm_graph_capture(efvm_ctx* ctx, int n_steps);
efvm_status efvm_graph_launch(efvm_ctx* ctx);

/* GhostCellIbmPlugin + CylinderIbmPlugin (crates/cfd_ibm/src/lib.rs:621-680):
 * a static z-aligned cylinder; `freestream` seeds cells that flip solid/fluid. */
efvm_status efvm_ibm_add_cylinder(efvm_ctx* ctx, double cx, double cy, double r, const double freestream[5]);
/* MarkerForce (pressure xyz, viscous xyz), FP64 (crates/cfd_ibm/src/lib.rs:442-480,601-615). */
efvm_status efvm_marker_force(efvm_ctx* ctx, double out[6]);
/* WallForce (pressure xyz, viscous xyz), FP64: the staircase estimate over solid-fluid faces
 * (crates/cfd_ibm/src/lib.rs:309-362, refreshed by wall_force_system, lib.rs:383-395), summed
 * in the reference's cell and face order.  Ghost-cell neighbours take the values the last
 * substage's boundary fill gave them; zero before the first substage creates the mask. */
efvm_status efvm_wall_force(efvm_ctx* ctx, double out[6]);
/* FP64 volume-weighted interior sums of (rho, rho_u, rho_v, rho_w, rho_E). */
efvm_status efvm_conservation_sums(efvm_ctx* ctx, double out[5]);
/* The same sums with no allocation and no host synchronisation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture, serving loops).
 * planes: the four stored components (rho, rho_u, rho_v, rho_E) of n_cells cells in caller
 * device memory, FP32 or FP64 (precision = EFVM_FP32 / EFVM_FP64), e.g. efvm_device_state's;
 * out5: 5 doubles of caller device memory, written (rho_w = 0) on `stream` (a hipStream_t,
 * null for the default stream).  The per-block partials live in a fixed 32 KiB device array
 * per GPU: calls that overlap on one GPU race on it.  1 <= n_cells <= EFVM_NOALLOC_MAX_CELLS,
 * else EFVM_ERR_ARG. */
#define EFVM_NOALLOC_MAX_CELLS 2147483647LL
efvm_status efvm_state_sums_noalloc(const void* const planes[4], int precision, long long n_cells, double cell_volume,
                                    double* out5, void* stream);

/* Raw device planes (rho, rho_u, rho_v, rho_E), nx*ny elements each of the
 * context precision (double or float), for zero-copy consumers. */
efvm_status efvm_device_state(efvm_ctx* ctx, void* planes[4]);
/* Busy time of the last efvm_step/efvm_graph_launch as measured by HIP events
 * around the stage kernels on the context stream, in milliseconds. */
efvm_status efvm_last_elapsed_ms(efvm_ctx* ctx, double* ms);

/* ---------------------------------------------------------------------------
 * Batched ensembles: many independent 2-D cases of uneven size advanced by one
 * stage launch per RK stage (a wave per case strip), each with its own gamma, boundaries and
 * CFL-limited dt.  The reference runs one App per case; this has no single
 * counterpart and replaces N copies of the context path above.  FP32, RK3.
 * ------------------------------------------------------------------------- */
typedef struct {
    int nx, ny;
    double lo_x, lo_y, hi_x, hi_y, dz;
    double gamma;
    efvm_bc sides[4];
} efvm_case;
typedef struct efvm_batch efvm_batch;

efvm_status efvm_batch_create(int n_cases, const efvm_case* cases, const efvm_scheme* scheme, efvm_batch** out);
void efvm_batch_destroy(efvm_batch* b);
const char* efvm_batch_last_error(const efvm_batch* b);
/* rows per wave strip for one case (case_id < 0: all cases); used by the autotuner */
efvm_status efvm_batch_set_launch(efvm_batch* b, int case_id, int strip_rows);
/* rows per wave strip for every case at once (n_cases ints) */
efvm_status efvm_batch_set_rows(efvm_batch* b, const int* rows);
/* launch grid: workgroups (0 = one wave per work item, the default; fewer make the
 * waves stride through the work table) and waves per workgroup */
/* Allow (1, the default) or forbid (0) the two-column strip kernel for the cases it takes
 * (periodic, outflow, slip, no-slip and symmetry sides, no features; nx even on a periodic x
 * axis).  No counterpart in the reference: a launch choice, results are bitwise the same. */
efvm_status efvm_batch_set_pair_kernel(efvm_batch* batch, int allow);
/* Allow (1, the default) or forbid (0) running stages 2 and 3 of the periodic two-column cases
 * (both axes periodic, nx even, ny >= 8) as one fused launch, a (stage 2, stage 3) wave pair per
 * strip of strip_rows rows (0 keeps the current height).  A fused case's U^n alternates between
 * two plane sets by step parity; upload, download, the CFL scan and snapshots follow it.  No
 * counterpart in the reference: a launch choice, results are bitwise the same. */
efvm_status efvm_batch_set_fused(efvm_batch* batch, int allow, int strip_rows);
/* The fused setting, its strip height and how many cases run fused (any pointer may be null). */
efvm_status efvm_batch_get_fused(const efvm_batch* batch, int* allow, int* strip_rows, int* n_fused);
efvm_status efvm_batch_set_grid(efvm_batch* b, int blocks, int waves_per_block);
efvm_status efvm_batch_upload(efvm_batch* b, int case_id, const double* cons5, size_t n_cells);
efvm_status efvm_batch_download(efvm_batch* b, int case_id, double* cons5, size_t n_cells);
efvm_status efvm_batch_step(efvm_batch* b, int n_steps);
efvm_status efvm_batch_graph_capture(efvm_batch* b, int n_steps);
efvm_status efvm_batch_graph_launch(efvm_batch* b);
efvm_status efvm_batch_clock(efvm_batch* b, int case_id, double* dt, double* time, long* step);
efvm_status efvm_batch_synchronize(efvm_batch* b);
efvm_status efvm_batch_last_elapsed_ms(efvm_batch* b, double* ms);

/* ---------------------------------------------------------------------------
 * State I/O (the reference's restart files and VTK output).
 * ------------------------------------------------------------------------- */
/* CFDRST01 checkpoint, the format of examples/restart_determinism/main.rs:177-226
 * (write_checkpoint / read_checkpoint): "CFDRST01", dt, time (f64), step (u64), rk stage
 * (u8), then the columns u, rhs, u0, each a u64 cell count and 5 f64 per cell over the
 * reference's raw padded array ((nx+4)(ny+4)(1+4) cells, x fastest), little-endian.
 * Ghost cells are written as 0 (the Halo and Boundary phases refill them), rhs as 0.
 * Save works at any substage.  Load takes a step-boundary checkpoint of the same grid
 * and resumes it bit for bit: the next step's CFL of the loaded state gives the dt the
 * uninterrupted run takes (crates/cfd_solver/tests/restart.rs:86-105). */
efvm_status efvm_checkpoint_save(efvm_ctx* ctx, const char* path);
efvm_status efvm_checkpoint_load(efvm_ctx* ctx, const char* path);
/* VTK legacy RECTILINEAR_GRID output of the interior cells, write_vtk_to of
 * crates/cfd_output/src/lib.rs:36-113: density, pressure, temperature, velocity and Mach
 * in 10-digit scientific notation (non-finite values as 0).  Host only: it takes a state in
 * the wire order (efvm_download_state), so it needs no device. */
efvm_status efvm_write_vtk(const efvm_grid* grid, const efvm_gas* gas, const double* cons5, size_t n_cells,
                           const char* path);

/* Device-side snapshot of the state and clock(s): op 0 saves, op 1 restores.
 * Lets the autotuner time real steps and then put the simulation back. */
efvm_status efvm_snapshot(efvm_ctx* ctx, int op);
efvm_status efvm_batch_snapshot(efvm_batch* b, int op);

/* ---------------------------------------------------------------------------
 * Learned face reconstruction.  The reference has no learned part: these entry
 * points extend its reconstruction choice (MusclPlugin::with_limiter,
 * crates/cfd_solver/src/lib.rs:806-842) with a network that stands in for the
 * limited MUSCL slope on coarse periodic grids.
 *
 * Per face side the network reads the 5-cell stencil of primitives (rho, u_n,
 * u_t, p) oriented toward the face (20 features, normalised per variable) and
 * returns 4 normalised face offsets:
 *   y = (x[+1] - x[-1]) / 4 + phi * (L x + bL),
 *   phi = sigmoid(W3 gelu(W2 gelu(W1 x + b1) + b2) + b3)      (per variable):
 * a learned high-order correction gated by an MLP limiter, 20 -> 64 -> 64 -> 4
 * with tanh-form GELU; inference runs the MLP on bf16 matrix cores with FP32
 * accumulation.  A face whose either side fails
 * positivity or the curvature-relaxed TVD bound uses the solver's MUSCL (MC
 * limiter) on both sides.  Parameters are one FP32 vector of
 * EFVM_MLP_NPARAM values: W1[64][20], b1[64], W2[64][64], b2[64], W3[4][64],
 * b3[4], L[4][20], bL[4] (row-major).
 * ------------------------------------------------------------------------- */
#define EFVM_MLP_NPARAM 5848
typedef struct efvm_mlp efvm_mlp;

efvm_status efvm_mlp_create(efvm_mlp** out);
void efvm_mlp_destroy(efvm_mlp* m);
const char* efvm_mlp_last_error(const efvm_mlp* m);
/* parameters in / out (n = EFVM_MLP_NPARAM); load also resets the Adam state */
efvm_status efvm_mlp_load(efvm_mlp* m, const float* params, size_t n);
efvm_status efvm_mlp_save(efvm_mlp* m, float* params, size_t n);
/* Training samples (rows of 20 features, 4 target offsets) held in device memory.
 * add_snapshot box-filters a fine periodic context's current state by `factor`
 * and appends 4 samples per coarse cell (x/y faces, both sides); the targets are
 * the fine solution's line averages over the coarse faces. */
efvm_status efvm_mlp_add_snapshot(efvm_mlp* m, efvm_ctx* fine, int factor);
/* As add_snapshot, but the features come from a coarse context's own state (nx/factor
 * by ny/factor, advanced to the same time) and only the targets from the fine one:
 * samples of the states the network meets inside the solver. */
efvm_status efvm_mlp_add_snapshot_pair(efvm_mlp* m, efvm_ctx* fine, efvm_ctx* coarse, int factor);
efvm_status efvm_mlp_add_samples(efvm_mlp* m, const float* X, const float* Y, size_t n);
efvm_status efvm_mlp_get_samples(efvm_mlp* m, size_t first, size_t n, float* X, float* Y);
efvm_status efvm_mlp_num_samples(efvm_mlp* m, size_t* n);
efvm_status efvm_mlp_clear_samples(efvm_mlp* m);
/* n_steps Adam steps on minibatches of `batch` samples drawn from the held
 * samples (counter hash of seed and step), learning rate cosine from lr0 to lr1;
 * train_skip = 0 keeps the linear skip (L, bL) fixed; losses (optional, n_steps
 * doubles) receive each minibatch's loss. */
efvm_status efvm_mlp_train_step(efvm_mlp* m, int n_steps, int batch, float lr0, float lr1, unsigned long long seed,
                                int train_skip, double* losses);
/* Closed-form fit of the linear skip (L, bL) to the held samples: the weighted
 * normal equations of y* - base over every `stride`-th sample, accumulated in FP64
 * on the device and solved on the host with a ridge of ridge * trace / 21.
 * constrained = 1 keeps the offsets exact for polynomial profiles up to degree 3
 * (per variable, only the coefficient along (1, -4, -4, 1) on cells -2, -1, +1, +2
 * is fitted; the unconstrained optimum of smooth training data amplifies grid-scale
 * noise in the solver). */
efvm_status efvm_mlp_fit_linear(efvm_mlp* m, size_t stride, double ridge, int constrained);
/* Loss and gradient over n given samples with the training kernels (FP32).  The loss
 * is the mean of w (y - y*)^2 with w = min((sigma / (0.01 scale))^2, 100) per variable
 * (the squared face-value error), sigma / scale recovered from the centre feature. */
efvm_status efvm_mlp_gradient(efvm_mlp* m, const float* X, const float* Y, size_t n, float* grad, double* loss);
/* Matrix-core inference of y on n given feature rows. */
efvm_status efvm_mlp_infer(efvm_mlp* m, const float* X, size_t n, float* Y);
/* Attach a copy of the network's current weights to an FP32 periodic inviscid
 * context (m = NULL detaches): every stage then builds its face states with it. */
efvm_status efvm_set_learned(efvm_ctx* ctx, efvm_mlp* m);
/* bypass = 1: the network's gate is evaluated only for face sides whose ungated
 * face value (phi = 1) fails the TVD bound with half its relaxation, i.e. where a
 * limiter can act; every other side keeps the correction fully applied.  0 (the
 * default): the network runs on every face side. */
efvm_status efvm_set_learned_bypass(efvm_ctx* ctx, int bypass);
/* Bypass stages with the 4-tap correction (the fitted network's form) run as one fused
 * stage kernel: every face's open-gate value, the strict-bound test, the Riemann flux and
 * the cell update in one pass, the fluxes kept on chip; cells next to a face that fails the
 * strict bound are finished by a second pass through the network.  The results are bitwise
 * those of the separate passes (allow = 0: light pass, deferred-face pass, flux-plane
 * stage).  Default: allowed.  No counterpart in the reference (which has no learned
 * reconstruction). */
efvm_status efvm_set_learned_fused(efvm_ctx* ctx, int allow);
/* allow, and whether the next stage takes the fused path */
efvm_status efvm_get_learned_fused(const efvm_ctx* ctx, int* allow, int* fits);
/* out[0] = faces that fell back to MUSCL, out[1] = faces evaluated, out[2] = face
 * sides whose gate the network evaluated (since reset) */
efvm_status efvm_mlp_infer_stats(efvm_ctx* ctx, int reset, unsigned long long out[3]);
/* Face primitives of the current state: 16 planes of nx*ny floats, plane
 * (axis*2 + side)*4 + var, side 0 = left / lower, var = rho, u, v, p, for the
 * east (axis 0) and north (axis 1) face of each cell. */
efvm_status efvm_learned_faces(efvm_ctx* ctx, float* out, size_t n);

/* ---------------------------------------------------------------------------
 * 3-D structured meshes: UniformMesh {nx, ny, nz, ng = 2, bounds_lo[3], bounds_hi[3]}
 * with any nz (field_core UniformMeshConfig; the solver systems of
 * crates/cfd_solver/src/lib.rs:93-122,170-339,449-677 over all three axes).  The same
 * plugins as efvm_ctx, six boundary sides, five stored components (rho_w included),
 * and the sphere body of crates/cfd_ibm/src/sphere.rs.  FP64 reproduces the reference's
 * operation order; FP32 runs the same operations in single precision.  Reconstruction
 * is first order or MUSCL; the learned reconstruction is 2-D only.
 * ------------------------------------------------------------------------- */
typedef struct { int nx, ny, nz; double lo[3], hi[3]; } efvm_grid3;
typedef struct efvm_ctx3 efvm_ctx3;

/* FieldDefaultPlugins{UniformMeshConfig} + CfdStatePlugin + IdealGasPlugin + SolverPlugin +
 * FluxPlugin / MusclPlugin + IntegratorPlugin for a 3-D mesh; every side starts as
 * SupersonicOutflow, the registry default (crates/cfd_boundary/src/lib.rs:366-381). */
efvm_status efvm3_create(const efvm_grid3* grid, const efvm_gas* gas, const efvm_scheme* scheme, efvm_ctx3** out);
void efvm3_destroy(efvm_ctx3* ctx);
const char* efvm3_last_error(const efvm_ctx3* ctx);
/* BoundaryRegistry (crates/cfd_boundary/src/lib.rs:369-449): sides[6] = XLo, XHi, YLo, YHi,
 * ZLo, ZHi; a half-periodic axis is EFVM_ERR_CONFIG. */
efvm_status efvm3_set_boundaries(efvm_ctx3* ctx, const efvm_bc sides[6]);
/* Stretched y (axis 1) or z (axis 2) spacing, UniformMeshConfig { y_edges, z_edges } of the
 * reference's mesh: the axis's n+1 strictly increasing cell edges from lo to hi.  Cell i spans
 * [e_i, e_{i+1}]: spacing e_{i+1} - e_i, centroid (e_i + e_{i+1}) / 2.  The face area over cell
 * volume of the flux divergence (crates/cfd_solver/src/lib.rs:418-434), the CFL spacing
 * (lib.rs:644-660), the viscous y / z weights from centroid distances (lib.rs:483-501) and the
 * conservation sums follow the cell.  Ghost centroids: periodic images on a periodic axis, else
 * mirror images across the boundary edge.  x stays uniform (the viscous x term takes
 * spacing(0, 0) for every cell, lib.rs:479-480); immersed bodies need uniform spacing.  At a
 * step boundary. */
efvm_status efvm3_set_edges(efvm_ctx3* ctx, int axis, const double* edges, size_t n_edges);
/* BoundaryInputs (crates/cfd_boundary/src/lib.rs:106-111) of a FromOutside side (kind
 * EFVM_BC_TABLE, lib.rs:115-147): cons5[5 n] ghost states in Component wire order and
 * given[n]; a face with given = 0 copies its interior cell (SupersonicOutflow).  One entry
 * per boundary face, the lower-numbered in-plane axis fastest: n = ny nz for an x side
 * (j + ny k), nx nz for a y side (i + nx k), nx ny for a z side (i + nx j).  May be called
 * again between steps; a captured graph reads the table in place once it exists. */
efvm_status efvm3_set_ghost_table(efvm_ctx3* ctx, int side, const double* cons5, const unsigned char* given, size_t n);
/* SolverConfig {cfl, fixed_dt} (crates/cfd_solver/src/lib.rs:39-58), at a step boundary. */
efvm_status efvm3_set_time_step(efvm_ctx3* ctx, double cfl, double fixed_dt);
/* Uniform x body force (crates/cfd_solver/tests/poiseuille_channel_validation.rs:105-120). */
efvm_status efvm3_set_body_force(efvm_ctx3* ctx, double gx);
/* outlet_sponge_system of examples/sphere_drag_validation/main.rs:119-160: after each
 * completed step, u += min(sigma s^2 dt, 1) (target - u) for cell centres x >= x_start,
 * s = (x - x_start) / (x_end - x_start) clamped to [0, 1]; target = conserved state. */
efvm_status efvm3_set_sponge(efvm_ctx3* ctx, double x_start, double x_end, double sigma, const double target[5]);
/* CfdState.u in Component wire order (5 doubles per cell, x fastest).  n_cells =
 * nx*ny*nz (interior; ghost cells are set to their nearest interior cell) or
 * (nx+4)(ny+4)(nz+4) (the reference's raw padded array, ghosts as given).
 * Upload resets the substage counter; both are synchronous. */
efvm_status efvm3_upload_state(efvm_ctx3* ctx, const double* cons5, size_t n_cells);
efvm_status efvm3_download_state(efvm_ctx3* ctx, double* cons5, size_t n_cells);
/* SolverState {dt, time, step, rk_stage} (crates/cfd_solver/src/lib.rs:61-73). */
efvm_status efvm3_get_clock(efvm_ctx3* ctx, double* dt, double* time, long* step, int* rk_stage);
efvm_status efvm3_set_clock(efvm_ctx3* ctx, double time, long step);
/* cfl_system (crates/cfd_solver/src/lib.rs:626-677) of the current state, FP64. */
efvm_status efvm3_cfl(efvm_ctx3* ctx, double* dt_out);
/* One app.run() (one substage through every phase) / n whole steps, enqueued only. */
efvm_status efvm3_run(efvm_ctx3* ctx);
efvm_status efvm3_step(efvm_ctx3* ctx, int n_steps);
efvm_status efvm3_synchronize(efvm_ctx3* ctx);
/* hipGraph of n_steps whole steps (even: the CFL slots alternate by step parity). */
efvm_status efvm3_graph_capture(efvm_ctx3* ctx, int n_steps);
efvm_status efvm3_graph_launch(efvm_ctx3* ctx);
/* SphereIbmPlugin::new(vec![SphereBody::fixed(center, r)], freestream).with_markers(n)
 * (crates/cfd_ibm/src/sphere.rs:82-253) with GhostCellIbmPlugin: solid mask, ghost list
 * with image points, and n_markers equal-area spherical-Fibonacci surface markers.
 * Static bodies, up to 8; all bodies of one context are spheres or all cylinders. */
efvm_status efvm3_ibm_add_sphere(efvm_ctx3* ctx, double cx, double cy, double cz, double r,
                                 const double freestream[5], int n_markers);
/* CylinderIbmPlugin (crates/cfd_ibm/src/lib.rs:155-254,566-598): z-aligned cylinder through
 * every layer, 64 x nz markers. */
efvm_status efvm3_ibm_add_cylinder(efvm_ctx3* ctx, double cx, double cy, double r, const double freestream[5]);
/* MarkerForce (pressure xyz, viscous xyz), FP64 (crates/cfd_ibm/src/lib.rs:442-480,601-615). */
efvm_status efvm3_marker_force(efvm_ctx3* ctx, double out[6]);
/* WallForce (pressure xyz, viscous xyz), FP64 staircase estimate over solid-fluid faces of the
 * raw state, ghost cells included (crates/cfd_ibm/src/lib.rs:309-362), in the reference's
 * summation order; zero before the first substage creates the mask. */
efvm_status efvm3_wall_force(efvm_ctx3* ctx, double out[6]);
/* FP64 volume-weighted interior sums of (rho, rho_u, rho_v, rho_w, rho_E). */
efvm_status efvm3_conservation_sums(efvm_ctx3* ctx, double out[5]);
/* Busy time of the last efvm3_step / efvm3_graph_launch (HIP events), milliseconds. */
efvm_status efvm3_last_elapsed_ms(efvm_ctx3* ctx, double* ms);
/* CFDRST01 checkpoints of examples/restart_determinism/main.rs:177-226 over the 3-D raw padded
 * array ((nx+4)(ny+4)(nz+4) cells): u with its ghost cells, rhs as 0, u0 the RK backup.  Load
 * takes a step-boundary file of the same grid and continues bit for bit
 * (crates/cfd_solver/tests/restart.rs:86-105). */
efvm_status efvm3_checkpoint_save(efvm_ctx3* ctx, const char* path);
efvm_status efvm3_checkpoint_load(efvm_ctx3* ctx, const char* path);
/* VTK legacy RECTILINEAR_GRID of nx*ny*nz interior cells, write_vtk_to of
 * crates/cfd_output/src/lib.rs:36-113 (the 2-D writer's fields and format).  Host only. */
efvm_status efvm3_write_vtk(const efvm_grid3* grid, const efvm_gas* gas, const double* cons5, size_t n_cells,
                            const char* path);
/* The same with stretched y / z axes (edges as efvm3_set_edges takes them, null = uniform):
 * each axis's coordinates are centroid - spacing/2 per cell and centroid + spacing/2 of the
 * last (crates/cfd_output/src/lib.rs:50-72). */
efvm_status efvm3_write_vtk_edges(const efvm_grid3* grid, const double* y_edges, const double* z_edges,
                                  const efvm_gas* gas, const double* cons5, size_t n_cells, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* EFVM_H */
