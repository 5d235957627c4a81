/*
 * efvm_oracle.c -- CPU FP64 restatement of the reference's explicit
 * compressible finite-volume update.  TEST INFRASTRUCTURE ONLY: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark use it as the checker.
 * The product path never links, loads or calls it.
 *
 * It follows the reference (Rust, SueHeir/dev_field_efvm) function by function,
 * keeping its data layout (array-of-structs ConsVar over a padded raw array
 * with ng = 2 ghost layers on every axis, nz = 1 for "2-D" cases), its face
 * sweep with a two-sided scatter, and its floating-point operation order.
 * Build with -ffp-contract=off: Rust does not contract multiply-adds.
 *
 *   ConsVar / PrimVar / Component wire order   crates/cfd_state/src/lib.rs:18-24,54-76,152-159
 *   IdealGas EOS, Sutherland, conductivity     crates/cfd_eos/src/lib.rs:43-87,108-117
 *   physical flux, Rusanov, HLLC(+Thornber)    crates/cfd_riemann/src/lib.rs:22-43,222-240,395-506,619-644
 *   flux_normal axis dispatch                  crates/cfd_riemann/src/lib.rs:76-93
 *   limiters, slopes, reconstruction, sweeps   crates/cfd_solver/src/lib.rs:93-122,170-339
 *   viscous terms                              crates/cfd_solver/src/lib.rs:449-560
 *   CFL                                        crates/cfd_solver/src/lib.rs:626-677
 *   Euler / SSP-RK3 substeps                   crates/cfd_solver/src/integrate.rs:46-101
 *   boundary conditions and the ghost fill     crates/cfd_boundary/src/lib.rs:173-362,574-601
 *   ghost-cell IBM, markers, staircase forces  crates/cfd_ibm/src/lib.rs:106-301,309-362,442-480,566-615,663-673
 *   parabolic inflow, outlet sponge            examples/cylinder_drag_validation/main.rs:151-229
 *
 * What the reference's own sources cannot fix (it delegates them to the
 * un-vendored field_core crate) is fixed here, and the GPU path follows the
 * same choices:
 *   - raw index = ir + NX*(jr + NY*kr), NX = nx+4, NY = ny+4 (x fastest);
 *   - cell centroid = lo + (i + 0.5)*h with i = raw - ng (ghosts included);
 *   - cell volume = (dx*dy)*dz; face area = dy*dz, dx*dz, dx*dy;
 *   - face sweep order = the structured fast path's order
 *     (crates/cfd_solver/src/lib.rs:410-446): cells k, j, i; per cell axes
 *     x, y, z; per axis the + face then, at the low edge, the - face;
 *   - axis neighbours of an edge cell are its ghost cells;
 *   - a periodic axis is a halo link: both ghost layers are copied from the
 *     opposite slab (crates/cfd_boundary/src/lib.rs:805-828), seam faces have
 *     no patch, and the ghost side of a seam face is not reconstructed
 *     (seam_mode 0, crates/cfd_solver/src/lib.rs:327-331); seam_mode 1
 *     reconstructs it from the wrapped cell, which makes the seam conservative;
 *   - locate(): s = (p - lo)/h - 0.5, i0 = floor(s), t = s - i0.
 * These are parity-unpinned by any reference test except where the tests
 * under tests/ say otherwise.  The fixed-problem digest of
 * examples/restart_determinism/main.rs:79 (tests/test_oracle3d.py) fixes the
 * sweep's axis order: the oracle matches it bit for bit, and with the axes
 * swept z, y, x it does not.
 *
 * orc_set_threads(s, n) runs the loops whose result does not depend on their
 * order (slopes, halo and boundary fills, viscous terms, RK update, CFL maximum)
 * on n OpenMP threads, and the face sweep in two phases: every face flux first,
 * then each cell's right-hand side summed in exactly the order the serial sweep
 * scatters into it.  The result is bitwise the serial one (tests/test_oracle.py
 * checks it); the serial form stays the definition.
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NG 2

typedef struct { double r, ru, rv, rw, re; } cv_t;           /* ConsVar */
typedef struct { double r, u, v, w, p, t; } pv_t;             /* PrimVar */

/* ---- enums shared with oracle/orc.py ---- */
enum { FLUX_RUSANOV = 0, FLUX_HLLC = 1, FLUX_HLLC_TEXTBOOK = 2 };
enum { RECON_NONE = 0, RECON_MUSCL = 1 };
enum { LIM_MINMOD = 0, LIM_MC = 1 };
enum { INTEG_EULER = 0, INTEG_RK3 = 1 };
enum { VISC_SUTHERLAND = 0, VISC_CONSTANT = 1 };
enum { BC_OUTFLOW = 0, BC_PERIODIC = 1, BC_SLIP = 2, BC_NOSLIP = 3, BC_SYMMETRY = 4,
       BC_SUP_INFLOW = 5, BC_SUB_INFLOW = 6, BC_PRESSURE_OUTLET = 7, BC_TABLE = 8 };
enum { STAGE_FIRST = 0, STAGE_SECOND = 1, STAGE_THIRD = 2 };

typedef struct {
    int nx, ny, nz;
    double lo[3], hi[3];
    double gamma, r_gas;
    int flux, recon, limiter, integrator;
    int viscous, visc_model;
    double mu_const;
    double cfl, fixed_dt;          /* fixed_dt <= 0: CFL-limited */
    int bc[6];                     /* XLo XHi YLo YHi ZLo ZHi */
    double bc_par[6][5];           /* inflow (rho,u,v,w,p|T), outlet p */
    int seam_mode;
    /* case systems of examples/cylinder_drag_validation/main.rs */
    int parabolic_inflow; double inflow_umax, inflow_h, inflow_rho;
    int sponge; double sp_x0, sp_x1, sp_sigma, sp_umax, sp_h, sp_rho, sp_p;
    int body_force; double body_gx;   /* crates/cfd_solver/tests/poiseuille_channel_validation.rs:105-120 */
    int n_cyl; double cyl[4][3];      /* cx, cy, r */
    double freestream[5];
    /* SphereIbmPlugin (crates/cfd_ibm/src/sphere.rs): fixed spheres, n Fibonacci markers */
    int n_sph; double sph[4][4];      /* cx, cy, cz, r */
    int sph_markers;
    /* uniform-target sponge of examples/sphere_drag_validation/main.rs:119-160 */
    int sp_uniform; double sp_target[5];
} orc_cfg;

typedef struct {
    orc_cfg c;
    int NX, NY, NZ;
    long ntot;
    double h[3];
    /* stretched y / z axes (UniformMeshConfig y_edges / z_edges): widths of the interior cells
       and centroids of every raw position along the axis; null: uniform spacing h */
    double *width[3], *cen[3];
    cv_t *u, *rhs, *u0;
    cv_t (*slope)[3];
    pv_t *prim;
    unsigned char *solid;
    double dt, time;
    long step;
    int stage;
    /* IBM */
    long n_ghost; long *ghost_idx; double (*ghost_img)[3];
    long n_mark; double (*mark)[7];  /* pos[3], normal[3], area */
    double marker_force[6], wall_force[6];
    int ibm_ready;
    /* FromOutside + BoundaryInputs (cfd_boundary:108-147): per boundary face of a side,
       5 conserved values and a given flag; not given -> SupersonicOutflow */
    double *tab[6]; long tab_n[6];
    /* threads > 1: the order-independent loops run in parallel and the face sweep in two
       phases (fluxes, then each cell's sum in the sweep's order): bitwise the serial result */
    int threads;
    cv_t *fplus[3];                  /* + face flux of every interior cell, per axis */
    cv_t *fminus[3];                 /* - face flux of the low-edge cells, per axis */
} orc_sim;

static inline long IDX(const orc_sim *s, int i, int j, int k) { return i + (long)s->NX * (j + (long)s->NY * k); }

/* ---------------- cfd_state / cfd_eos ---------------- */
static inline cv_t cv_add(cv_t a, cv_t b) { cv_t o = {a.r + b.r, a.ru + b.ru, a.rv + b.rv, a.rw + b.rw, a.re + b.re}; return o; }
static inline cv_t cv_sub(cv_t a, cv_t b) { cv_t o = {a.r - b.r, a.ru - b.ru, a.rv - b.rv, a.rw - b.rw, a.re - b.re}; return o; }
static inline cv_t cv_mul(cv_t a, double s) { cv_t o = {a.r * s, a.ru * s, a.rv * s, a.rw * s, a.re * s}; return o; }
static inline void vel_of(const cv_t *u, double v[3]) {            /* ConsVar::velocity, cfd_state:40-43 */
    double inv = 1.0 / u->r;
    v[0] = u->ru * inv; v[1] = u->rv * inv; v[2] = u->rw * inv;
}
static inline double kinetic(const cv_t *u) {                      /* cfd_state:46-49 */
    return 0.5 * (u->ru * u->ru + u->rv * u->rv + u->rw * u->rw) / u->r;
}
static inline double pressure(const cv_t *u, double g) { return (g - 1.0) * (u->re - kinetic(u)); } /* cfd_eos:61-63 */
static inline double sound(const cv_t *u, double g) { double p = pressure(u, g); return sqrt(g * p / u->r); } /* :65-68 */
static inline pv_t cons_to_prim(const cv_t *u, double g, double R) {  /* cfd_eos:44-53 */
    double inv = 1.0 / u->r;
    double vu = u->ru * inv, vv = u->rv * inv, vw = u->rw * inv;
    double ke = 0.5 * u->r * (vu * vu + vv * vv + vw * vw);
    double p = (g - 1.0) * (u->re - ke);
    double t = p * inv / R;
    pv_t o = {u->r, vu, vv, vw, p, t};
    return o;
}
static inline cv_t prim_to_cons(const pv_t *p, double g) {          /* cfd_eos:55-59 */
    double ke = 0.5 * p->r * (p->u * p->u + p->v * p->v + p->w * p->w);
    double re = p->p / (g - 1.0) + ke;
    cv_t o = {p->r, p->r * p->u, p->r * p->v, p->r * p->w, re};
    return o;
}
double orc_sutherland(double t) {                                   /* cfd_eos:75-80 */
    const double MU_REF = 1.827e-5, T_REF = 291.15, S = 120.0;
    return MU_REF * pow(t / T_REF, 1.5) * (T_REF + S) / (t + S);
}
static inline double visc_mu(const orc_cfg *c, double t) {           /* cfd_eos:111-116 */
    return c->visc_model == VISC_CONSTANT ? c->mu_const : orc_sutherland(t);
}
static inline double conductivity(double mu, double g, double R) {  /* cfd_eos:83-87 */
    double cp = g * R / (g - 1.0);
    return mu * cp / 0.72;
}

/* ---------------- cfd_riemann ---------------- */
static inline cv_t phys_flux(const cv_t *u, double g, int dir) {     /* cfd_riemann:22-43 */
    double p = pressure(u, g);
    double v[3]; vel_of(u, v);
    double vn = v[dir];
    cv_t f = {u->r * vn, u->ru * vn, u->rv * vn, u->rw * vn, (u->re + p) * vn};
    if (dir == 0) f.ru += p; else if (dir == 1) f.rv += p; else f.rw += p;
    return f;
}
static cv_t rusanov(const cv_t *L, const cv_t *R, double g, int dir) { /* cfd_riemann:222-240 */
    cv_t fl = phys_flux(L, g, dir), fr = phys_flux(R, g, dir);
    double vl[3], vr[3]; vel_of(L, vl); vel_of(R, vr);
    double al = sound(L, g), ar = sound(R, g);
    double lam = fmax(fabs(vl[dir]) + al, fabs(vr[dir]) + ar);
    return cv_sub(cv_mul(cv_add(fl, fr), 0.5), cv_mul(cv_sub(*R, *L), 0.5 * lam));
}
static cv_t star_state(const cv_t *u, const double vel[3], double p, double un, double s, double ss, int dir) {
    /* cfd_riemann:619-644 */
    double rs = u->r * (s - un) / (s - ss);
    double epm = u->re / u->r;
    double es = epm + (ss - un) * (ss + p / (u->r * (s - un)));
    cv_t o;
    o.r = rs;
    o.ru = dir == 0 ? rs * ss : rs * vel[0];
    o.rv = dir == 1 ? rs * ss : rs * vel[1];
    o.rw = dir == 2 ? rs * ss : rs * vel[2];
    o.re = rs * es;
    return o;
}
static void thornber(const cv_t *L, const cv_t *R, double g, cv_t *Lc, cv_t *Rc, double *al, double *ar) {
    /* cfd_riemann:456-506 */
    double vl[3], vr[3]; vel_of(L, vl); vel_of(R, vr);
    double vlsq = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2];
    double vrsq = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
    double a_l = sound(L, g), a_r = sound(R, g);
    double vmax = fmax(vlsq, vrsq), amax = fmax(a_l, a_r);
    double z = (vmax >= amax * amax) ? 1.0 : sqrt(vmax) / amax;
    double avg[3], dv[3], vlc[3], vrc[3];
    for (int d = 0; d < 3; d++) {
        avg[d] = 0.5 * (vl[d] + vr[d]);
        dv[d] = 0.5 * z * (vl[d] - vr[d]);
        vlc[d] = avg[d] + dv[d];
        vrc[d] = avg[d] - dv[d];
    }
    double vlcsq = vlc[0] * vlc[0] + vlc[1] * vlc[1] + vlc[2] * vlc[2];
    double vrcsq = vrc[0] * vrc[0] + vrc[1] * vrc[1] + vrc[2] * vrc[2];
    Lc->r = L->r; Lc->ru = L->r * vlc[0]; Lc->rv = L->r * vlc[1]; Lc->rw = L->r * vlc[2];
    Lc->re = L->re + 0.5 * L->r * (vlcsq - vlsq);
    Rc->r = R->r; Rc->ru = R->r * vrc[0]; Rc->rv = R->r * vrc[1]; Rc->rw = R->r * vrc[2];
    Rc->re = R->re + 0.5 * R->r * (vrcsq - vrsq);
    *al = a_l; *ar = a_r;
}
static cv_t hllc(const cv_t *L, const cv_t *R, double g, int dir, int use_thornber) {
    /* Hllc::flux_with_thornber, cfd_riemann:395-440 */
    cv_t lc, rc; double al, ar;
    if (use_thornber) thornber(L, R, g, &lc, &rc, &al, &ar);
    else { al = sound(L, g); ar = sound(R, g); lc = *L; rc = *R; }
    double vl[3], vr[3]; vel_of(&lc, vl); vel_of(&rc, vr);
    double pl = pressure(&lc, g), pr = pressure(&rc, g);
    double unl = vl[dir], unr = vr[dir];
    double sl = fmin(unl - al, unr - ar);
    double sr = fmax(unl + al, unr + ar);
    double fl = lc.r * (sl - unl);
    double fr = rc.r * (sr - unr);
    double ss = ((pr - pl) + fl * unl - fr * unr) / (fl - fr);
    if (sl >= 0.0) return phys_flux(&lc, g, dir);
    if (ss >= 0.0) {
        cv_t us = star_state(&lc, vl, pl, unl, sl, ss, dir);
        return cv_add(phys_flux(&lc, g, dir), cv_mul(cv_sub(us, lc), sl));
    }
    if (sr >= 0.0) {
        cv_t us = star_state(&rc, vr, pr, unr, sr, ss, dir);
        return cv_add(phys_flux(&rc, g, dir), cv_mul(cv_sub(us, rc), sr));
    }
    return phys_flux(&rc, g, dir);
}
static inline cv_t num_flux(int scheme, const cv_t *L, const cv_t *R, double g, int dir) {
    if (scheme == FLUX_RUSANOV) return rusanov(L, R, g, dir);
    return hllc(L, R, g, dir, scheme == FLUX_HLLC);
}
/* RiemannFlux::flux_normal for an axis-aligned unit normal (cfd_riemann:76-93). */
static inline cv_t flux_normal_axis(int scheme, const cv_t *own, const cv_t *oth, double g, int dir, int sign) {
    if (sign > 0) return num_flux(scheme, own, oth, g, dir);
    return cv_mul(num_flux(scheme, oth, own, g, dir), -1.0);
}

/* Exported scalar kernels for known-answer tests. */
void orc_flux(int scheme, const double *L, const double *R, double g, int dir, double *out) {
    cv_t l = {L[0], L[1], L[2], L[3], L[4]}, r = {R[0], R[1], R[2], R[3], R[4]};
    cv_t f = num_flux(scheme, &l, &r, g, dir);
    out[0] = f.r; out[1] = f.ru; out[2] = f.rv; out[3] = f.rw; out[4] = f.re;
}
void orc_flux_normal_axis(int scheme, const double *O, const double *T, double g, int dir, int sign, double *out) {
    cv_t o = {O[0], O[1], O[2], O[3], O[4]}, t = {T[0], T[1], T[2], T[3], T[4]};
    cv_t f = flux_normal_axis(scheme, &o, &t, g, dir, sign);
    out[0] = f.r; out[1] = f.ru; out[2] = f.rv; out[3] = f.rw; out[4] = f.re;
}
void orc_physical_flux(const double *U, double g, int dir, double *out) {
    cv_t u = {U[0], U[1], U[2], U[3], U[4]};
    cv_t f = phys_flux(&u, g, dir);
    out[0] = f.r; out[1] = f.ru; out[2] = f.rv; out[3] = f.rw; out[4] = f.re;
}
void orc_prim_to_cons(const double *P, double g, double *out) {
    pv_t p = {P[0], P[1], P[2], P[3], P[4], P[5]};
    cv_t c = prim_to_cons(&p, g);
    out[0] = c.r; out[1] = c.ru; out[2] = c.rv; out[3] = c.rw; out[4] = c.re;
}
void orc_cons_to_prim(const double *U, double g, double R, double *out) {
    cv_t u = {U[0], U[1], U[2], U[3], U[4]};
    pv_t p = cons_to_prim(&u, g, R);
    out[0] = p.r; out[1] = p.u; out[2] = p.v; out[3] = p.w; out[4] = p.p; out[5] = p.t;
}
double orc_sound_speed(const double *U, double g) { cv_t u = {U[0], U[1], U[2], U[3], U[4]}; return sound(&u, g); }

/* ---------------- limiters (cfd_solver:170-211) ---------------- */
static inline double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return fabs(a) < fabs(b) ? a : b;
}
static inline double mc(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    double s = a > 0.0 ? 1.0 : -1.0;
    return s * fmin(fmin(2.0 * fabs(a), 2.0 * fabs(b)), 0.5 * fabs(a + b));
}
double orc_limiter(int lim, double a, double b) { return lim == LIM_MC ? mc(a, b) : minmod(a, b); }
static inline cv_t limited(cv_t f, cv_t b, int lim) {
    cv_t o;
    if (lim == LIM_MC) { o.r = mc(f.r, b.r); o.ru = mc(f.ru, b.ru); o.rv = mc(f.rv, b.rv); o.rw = mc(f.rw, b.rw); o.re = mc(f.re, b.re); }
    else { o.r = minmod(f.r, b.r); o.ru = minmod(f.ru, b.ru); o.rv = minmod(f.rv, b.rv); o.rw = minmod(f.rw, b.rw); o.re = minmod(f.re, b.re); }
    return o;
}

/* ---------------- mesh helpers (field_core semantics, see header) ---------------- */
static inline double centroid(const orc_sim *s, int d, int raw) {
    return s->cen[d] ? s->cen[d][raw] : s->c.lo[d] + ((double)(raw - NG) + 0.5) * s->h[d];
}
/* StructuredMesh::spacing(axis, i) of interior cell i */
static inline double spacing(const orc_sim *s, int d, int i) { return s->width[d] ? s->width[d][i] : s->h[d]; }
static inline int is_local(const orc_sim *s, int i, int j, int k) {
    return i >= NG && i < NG + s->c.nx && j >= NG && j < NG + s->c.ny && k >= NG && k < NG + s->c.nz;
}
/* cell volume of interior cell (., j, k): dx dy_j dz_k (x is uniform) */
static inline double cell_volume_at(const orc_sim *s, int j, int k) { return s->h[0] * spacing(s, 1, j) * spacing(s, 2, k); }
static inline double cell_volume(const orc_sim *s) { return cell_volume_at(s, 0, 0); }
static inline int axis_periodic(const orc_sim *s, int d) { return s->c.bc[2 * d] == BC_PERIODIC && s->c.bc[2 * d + 1] == BC_PERIODIC; }

/* ---------------- lifecycle ---------------- */
orc_sim *orc_new(const orc_cfg *cfg) {
    for (int d = 0; d < 3; d++)
        if ((cfg->bc[2 * d] == BC_PERIODIC) != (cfg->bc[2 * d + 1] == BC_PERIODIC)) return NULL; /* cfd_boundary:437-448 */
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->c = *cfg;
    s->NX = cfg->nx + 2 * NG; s->NY = cfg->ny + 2 * NG; s->NZ = cfg->nz + 2 * NG;
    s->ntot = (long)s->NX * s->NY * s->NZ;
    int n[3] = {cfg->nx, cfg->ny, cfg->nz};
    for (int d = 0; d < 3; d++) s->h[d] = (cfg->hi[d] - cfg->lo[d]) / (double)n[d];
    s->u = (cv_t *)calloc(s->ntot, sizeof(cv_t));
    s->rhs = (cv_t *)calloc(s->ntot, sizeof(cv_t));
    s->u0 = (cv_t *)calloc(s->ntot, sizeof(cv_t));
    s->slope = calloc(s->ntot, sizeof(*s->slope));
    s->prim = (pv_t *)calloc(s->ntot, sizeof(pv_t));
    s->solid = (unsigned char *)calloc(s->ntot, 1);
    s->stage = STAGE_FIRST;
    return s;
}
void orc_free(orc_sim *s) {
    if (!s) return;
    free(s->u); free(s->rhs); free(s->u0); free(s->slope); free(s->prim); free(s->solid);
    for (int q = 0; q < 6; q++) free(s->tab[q]);
    free(s->ghost_idx); free(s->ghost_img); free(s->mark);
    for (int d = 0; d < 3; d++) { free(s->fplus[d]); free(s->fminus[d]); free(s->width[d]); free(s->cen[d]); }
    free(s);
}
long orc_ntot(const orc_sim *s) { return s->ntot; }
/* Number of OpenMP threads for the step (1 = the serial sweep); returns the count in use. */
int orc_set_threads(orc_sim *s, int n) {
#ifdef _OPENMP
    s->threads = n > 1 ? n : 1;
#else
    (void)n;
    s->threads = 1;
#endif
    return s->threads;
}
#define OMP_THREADS(s) num_threads((s)->threads > 1 ? (s)->threads : 1) if((s)->threads > 1)

/* Stretched axis (UniformMeshConfig y_edges / z_edges): n+1 increasing cell-edge coordinates
 * of axis d (1 = y, 2 = z) from lo to hi.  Interior cell i spans [e_i, e_{i+1}]: spacing
 * e_{i+1} - e_i, centroid (e_i + e_{i+1}) / 2.  Ghost centroids: on a periodic axis the
 * periodic image of the opposite cell, else the mirror image of the interior cell across the
 * boundary edge.  x stays uniform (the viscous x term uses spacing(0, 0) for every cell,
 * crates/cfd_solver/src/lib.rs:479-480), and immersed bodies need a uniform mesh.
 * Returns 0, or -1 for a bad axis, count, order, end points, or a body. */
static int n_bodies(const orc_sim *s);
int orc_set_edges(orc_sim *s, int d, const double *e, int n_edges) {
    int n = d == 1 ? s->c.ny : (d == 2 ? s->c.nz : -1);
    if (n < 0 || n_edges != n + 1 || n_bodies(s) > 0) return -1;
    if (e[0] != s->c.lo[d] || e[n] != s->c.hi[d]) return -1;
    for (int i = 0; i < n; i++) if (!(e[i + 1] > e[i])) return -1;
    free(s->width[d]); free(s->cen[d]);
    s->width[d] = (double *)malloc(sizeof(double) * n);
    s->cen[d] = (double *)malloc(sizeof(double) * (n + 2 * NG));
    for (int i = 0; i < n; i++) { s->width[d][i] = e[i + 1] - e[i]; s->cen[d][NG + i] = 0.5 * (e[i] + e[i + 1]); }
    double len = e[n] - e[0];
    for (int g = 1; g <= NG; g++) {
        double *lo = &s->cen[d][NG - g], *hi = &s->cen[d][NG + n - 1 + g];
        if (axis_periodic(s, d)) { *lo = s->cen[d][NG + n - g] - len; *hi = s->cen[d][NG + g - 1] + len; }
        else { *lo = 2.0 * e[0] - s->cen[d][NG + g - 1]; *hi = 2.0 * e[n] - s->cen[d][NG + n - g]; }
    }
    return 0;
}

/* Raw (padded) state, Component wire order (rho, rho_u, rho_v, rho_w, rho_E) per cell. */
void orc_set_raw(orc_sim *s, const double *src) { memcpy(s->u, src, sizeof(cv_t) * s->ntot); }
void orc_get_raw(const orc_sim *s, double *dst) { memcpy(dst, s->u, sizeof(cv_t) * s->ntot); }
/* The rhs and u0 columns over the raw array, as a CFDRST01 file holds them. */
void orc_get_rhs_raw(const orc_sim *s, double *dst) { memcpy(dst, s->rhs, sizeof(cv_t) * s->ntot); }
void orc_get_u0_raw(const orc_sim *s, double *dst) { memcpy(dst, s->u0, sizeof(cv_t) * s->ntot); }
/* Interior cells only, i fastest, then j, then k; 5 values per cell. */
void orc_set_interior(orc_sim *s, const double *src) {
    long q = 0;
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++, q++)
        memcpy(&s->u[IDX(s, i + NG, j + NG, k + NG)], src + 5 * q, sizeof(cv_t));
}
void orc_get_interior(const orc_sim *s, double *dst) {
    long q = 0;
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++, q++)
        memcpy(dst + 5 * q, &s->u[IDX(s, i + NG, j + NG, k + NG)], sizeof(cv_t));
}
void orc_info(const orc_sim *s, double *dt, double *time, long *step, int *stage) {
    *dt = s->dt; *time = s->time; *step = s->step; *stage = s->stage;
}
void orc_set_clock(orc_sim *s, double dt, double time, long step, int stage) { s->dt = dt; s->time = time; s->step = step; s->stage = stage; }

/* BoundaryInputs for a FromOutside side: n faces x (rho, rho_u, rho_v, rho_w, rho_E, given) */
int orc_set_ghost_table(orc_sim *s, int side, const double *tab6, long n) {
    if (side < 0 || side > 5 || s->c.bc[side] != BC_TABLE) return -1;
    free(s->tab[side]);
    s->tab[side] = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    if (!s->tab[side]) return -1;
    memcpy(s->tab[side], tab6, sizeof(double) * 6 * (size_t)n);
    s->tab_n[side] = n;
    return 0;
}

/* ---------------- initial conditions (evaluated at every raw cell, cfd_solver:681-686) ---------------- */
/* Shu isentropic vortex, examples/isentropic_vortex/main.rs:97-125; `periodic` selects the
 * minimum-image distance on a periodic [0,L]^2 box (the mean flow is then (mean_u, mean_v)). */
static cv_t vortex_at(double L, double beta, double gam, double mu_, double mv, double t, double x, double y, int periodic) {
    double cx = 0.5 * L + mu_ * t, cy = 0.5 * L + mv * t;
    double dx = x - cx, dy = y - cy;
    if (periodic) { dx = dx - L * round(dx / L); dy = dy - L * round(dy / L); }
    double r2 = dx * dx + dy * dy;
    double f = exp(0.5 * (1.0 - r2));
    const double PI = 3.14159265358979323846;
    double du = -beta / (2.0 * PI) * dy * f;
    double dv = beta / (2.0 * PI) * dx * f;
    double dtemp = -(gam - 1.0) * beta * beta / (8.0 * gam * (PI * PI)) * exp(1.0 - r2);
    double temp = 1.0 + dtemp;
    double rho = pow(temp, 1.0 / (gam - 1.0));
    double p = pow(rho, gam);
    pv_t pv = {rho, mu_ + du, mv + dv, 0.0, p, temp};
    return prim_to_cons(&pv, gam);
}
void orc_vortex_point(double L, double beta, double gam, double mu_, double mv, double t, double x, double y, int periodic, double *out) {
    cv_t c = vortex_at(L, beta, gam, mu_, mv, t, x, y, periodic);
    out[0] = c.r; out[1] = c.ru; out[2] = c.rv; out[3] = c.rw; out[4] = c.re;
}
void orc_ic_vortex(orc_sim *s, double L, double beta, double gam, double mu_, double mv, int periodic) {
    for (int k = 0; k < s->NZ; k++) for (int j = 0; j < s->NY; j++) for (int i = 0; i < s->NX; i++)
        s->u[IDX(s, i, j, k)] = vortex_at(L, beta, gam, mu_, mv, 0.0, centroid(s, 0, i), centroid(s, 1, j), periodic);
}
/* Volume-weighted L1 density error against the translated vortex (main.rs:189-211). */
double orc_vortex_l1(const orc_sim *s, double L, double beta, double gam, double mu_, double mv, int periodic) {
    double num = 0.0, vol = 0.0;
    for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++) {
        const double cvv = cell_volume_at(s, j, 0);
        long c = IDX(s, i + NG, j + NG, NG);
        double x = centroid(s, 0, i + NG), y = centroid(s, 1, j + NG);
        double ex;
        if (periodic) {
            /* exact solution on the periodic box: wrap the translated centre */
            double cx = fmod(0.5 * L + mu_ * s->time, L), cy = fmod(0.5 * L + mv * s->time, L);
            double dx = x - cx, dy = y - cy;
            dx = dx - L * round(dx / L); dy = dy - L * round(dy / L);
            cv_t e = vortex_at(L, beta, gam, 0.0, 0.0, 0.0, 0.5 * L + dx, 0.5 * L + dy, 0);
            ex = e.r;
        } else {
            cv_t e = vortex_at(L, beta, gam, mu_, mv, s->time, x, y, 0);
            ex = e.r;
        }
        num += fabs(s->u[c].r - ex) * cvv;
        vol += cvv;
    }
    return num / vol;
}
/* Periodic contact pulse, examples/periodic_advected_pulse/main.rs:62-73. */
static cv_t pulse_at(double L, double cen, double sig, double amp, double gam, double vel, double pr, double t, double x) {
    double c = fmod(cen + vel * t, L); if (c < 0) c += L;          /* rem_euclid */
    double d = x - c; d = d - L * round(d / L);
    double q = d / sig;
    double rho = 1.0 + amp * exp(-0.5 * (q * q));
    double temp = pr / rho;
    pv_t pv = {rho, vel, 0.0, 0.0, pr, temp};
    return prim_to_cons(&pv, gam);
}
void orc_ic_pulse(orc_sim *s, double L, double cen, double sig, double amp, double gam, double vel, double pr) {
    for (int k = 0; k < s->NZ; k++) for (int j = 0; j < s->NY; j++) for (int i = 0; i < s->NX; i++)
        s->u[IDX(s, i, j, k)] = pulse_at(L, cen, sig, amp, gam, vel, pr, 0.0, centroid(s, 0, i));
}
double orc_pulse_l1(const orc_sim *s, double L, double cen, double sig, double amp, double gam, double vel, double pr) {
    double num = 0.0, vol = 0.0, cvv = cell_volume(s);
    for (int i = 0; i < s->c.nx; i++) {
        long c = IDX(s, i + NG, NG, NG);
        cv_t e = pulse_at(L, cen, sig, amp, gam, vel, pr, s->time, centroid(s, 0, i + NG));
        num += fabs(s->u[c].r - e.r) * cvv; vol += cvv;
    }
    return num / vol;
}
/* Piecewise state: x < x0 ? A : B (Sod, crates/cfd_solver/tests/sod_exact_convergence.rs:201-206). */
void orc_ic_split_x(orc_sim *s, double x0, const double *A, const double *B) {
    cv_t a = {A[0], A[1], A[2], A[3], A[4]}, b = {B[0], B[1], B[2], B[3], B[4]};
    for (int k = 0; k < s->NZ; k++) for (int j = 0; j < s->NY; j++) for (int i = 0; i < s->NX; i++)
        s->u[IDX(s, i, j, k)] = centroid(s, 0, i) < x0 ? a : b;
}
void orc_ic_uniform(orc_sim *s, const double *A) {
    cv_t a = {A[0], A[1], A[2], A[3], A[4]};
    for (long c = 0; c < s->ntot; c++) s->u[c] = a;
}
/* Parabolic channel profile, examples/cylinder_drag_validation/main.rs:268-272. */
void orc_ic_channel(orc_sim *s, double um, double h, double rho0, double p0, double R) {
    for (int k = 0; k < s->NZ; k++) for (int j = 0; j < s->NY; j++) for (int i = 0; i < s->NX; i++) {
        double y = centroid(s, 1, j);
        double uy = 4.0 * um * y * (h - y) / (h * h);
        pv_t pv = {rho0, uy, 0.0, 0.0, p0, p0 / (rho0 * R)};
        s->u[IDX(s, i, j, k)] = prim_to_cons(&pv, s->c.gamma);
    }
}
/* Cell-volume-weighted interior sums of the 5 conserved components. */
void orc_sums(const orc_sim *s, double *out) {
    for (int q = 0; q < 5; q++) out[q] = 0.0;
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++) {
        const double cvv = cell_volume_at(s, j, k);
        const cv_t *u = &s->u[IDX(s, i + NG, j + NG, k + NG)];
        out[0] += u->r * cvv; out[1] += u->ru * cvv; out[2] += u->rv * cvv; out[3] += u->rw * cvv; out[4] += u->re * cvv;
    }
}

/* ---------------- systems ---------------- */
#pragma omp declare reduction(max_fmax : double : omp_out = fmax(omp_out, omp_in)) initializer(omp_priv = 0.0)
static void cfl_system(orc_sim *s) {                                /* cfd_solver:626-677 */
    if (s->stage != STAGE_FIRST) return;
    if (s->c.fixed_dt > 0.0) { s->dt = s->c.fixed_dt; return; }
    double g = s->c.gamma, lmax = 0.0;
    /* a maximum does not depend on the visiting order (fmax ignores a NaN lambda either way) */
#pragma omp parallel for OMP_THREADS(s) reduction(max_fmax : lmax)
    for (int i = 0; i < s->c.nx; i++) for (int j = 0; j < s->c.ny; j++) for (int k = 0; k < s->c.nz; k++) {
        const double dx = spacing(s, 0, i), dy = spacing(s, 1, j), dz = spacing(s, 2, k);
        const cv_t *u = &s->u[IDX(s, i + NG, j + NG, k + NG)];
        if (u->r <= 0.0) continue;
        double v[3]; vel_of(u, v);
        double a = sound(u, g);
        double lam = (fabs(v[0]) + a) / dx + (fabs(v[1]) + a) / dy + (fabs(v[2]) + a) / dz;
        if (s->c.viscous) {
            pv_t p = cons_to_prim(u, g, s->c.r_gas);
            double nu = visc_mu(&s->c, p.t) / u->r;
            lam += 2.0 * nu * (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
        }
        lmax = fmax(lmax, lam);
    }
    s->dt = lmax > 0.0 ? s->c.cfl / lmax : INFINITY;
}
double orc_cfl_dt(orc_sim *s) { int st = s->stage; s->stage = STAGE_FIRST; double keep = s->dt; cfl_system(s); double dt = s->dt; s->dt = keep; s->stage = st; return dt; }

static void halo_system(orc_sim *s) {                               /* periodic halo link */
    int n[3] = {s->c.nx, s->c.ny, s->c.nz};
    for (int d = 0; d < 3; d++) {
        if (!axis_periodic(s, d)) continue;
        /* within one axis every write is a ghost cell and every read an interior cell along d */
#pragma omp parallel for collapse(2) OMP_THREADS(s)
        for (int k = 0; k < s->NZ; k++) for (int j = 0; j < s->NY; j++) for (int i = 0; i < s->NX; i++) {
            int id[3] = {i, j, k};
            int r = id[d];
            if (r >= NG && r < NG + n[d]) continue;
            int src[3] = {i, j, k};
            src[d] = r < NG ? r + n[d] : r - n[d];
            /* only copy from interior along d; other axes as they are */
            s->u[IDX(s, i, j, k)] = s->u[IDX(s, src[0], src[1], src[2])];
        }
    }
}

static cv_t bc_ghost_of(const orc_cfg *c, int side, const cv_t *in) {  /* cfd_boundary:173-362 */
    double g = c->gamma;
    const double *par = c->bc_par[side];
    cv_t o = *in;
    int axis = side / 2;
    switch (c->bc[side]) {
    case BC_OUTFLOW: return o;
    case BC_SLIP: case BC_SYMMETRY:
        if (axis == 0) o.ru = -o.ru; else if (axis == 1) o.rv = -o.rv; else o.rw = -o.rw;
        return o;
    case BC_NOSLIP: o.ru = -o.ru; o.rv = -o.rv; o.rw = -o.rw; return o;
    case BC_SUP_INFLOW: { pv_t p = {par[0], par[1], par[2], par[3], par[4], par[4] / (par[0] * 287.058)}; return prim_to_cons(&p, g); }
    case BC_SUB_INFLOW: { double pin = pressure(in, g); pv_t p = {par[0], par[1], par[2], par[3], pin, par[4]}; return prim_to_cons(&p, g); }
    case BC_PRESSURE_OUTLET: { pv_t p = cons_to_prim(in, g, c->r_gas); p.p = par[0]; p.t = par[0] / (p.r * 287.058); return prim_to_cons(&p, g); }
    default: return o;
    }
}
static cv_t bc_ghost(const orc_sim *s, int side, const cv_t *in) { return bc_ghost_of(&s->c, side, in); }
/* BoundaryCondition::ghost of one stock condition for a box side (side 0..5 = XLo..ZHi),
 * par = its parameters (inflow rho, u, v, w, p | T; outlet p): the unit-test entry point. */
void orc_bc_ghost(int kind, int side, const double *par, const double *interior, double gamma, double r_gas, double *out) {
    orc_cfg c;
    memset(&c, 0, sizeof(c));
    c.gamma = gamma; c.r_gas = r_gas;
    c.bc[side] = kind;
    for (int q = 0; q < 5; q++) c.bc_par[side][q] = par ? par[q] : 0.0;
    cv_t in = {interior[0], interior[1], interior[2], interior[3], interior[4]};
    cv_t o = bc_ghost_of(&c, side, &in);
    out[0] = o.r; out[1] = o.ru; out[2] = o.rv; out[3] = o.rw; out[4] = o.re;
}

static void boundary_system(orc_sim *s) {                           /* cfd_boundary:574-601 */
    int nx = s->c.nx, ny = s->c.ny, nz = s->c.nz;
    for (int side = 0; side < 6; side++) {
        if (s->c.bc[side] == BC_PERIODIC) continue;
        int axis = side / 2, hi = side & 1;
        int n[3] = {nx, ny, nz};
        int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
#pragma omp parallel for collapse(2) OMP_THREADS(s)
        for (int q2 = 0; q2 < n[a2]; q2++) for (int q1 = 0; q1 < n[a1]; q1++) {
            int own[3], gh[3];
            own[a1] = q1 + NG; own[a2] = q2 + NG; gh[a1] = own[a1]; gh[a2] = own[a2];
            own[axis] = hi ? NG + n[axis] - 1 : NG;
            gh[axis] = hi ? NG + n[axis] : NG - 1;
            cv_t in = s->u[IDX(s, own[0], own[1], own[2])];
            if (s->c.bc[side] == BC_TABLE) {
                /* the face index over the side, the lower-numbered in-plane axis fastest:
                 * x side j + ny k, y side i + nx k, z side i + nx j (nz = 1: the row of an
                 * x side, the column of a y side) */
                long f = axis == 0 ? q1 + (long)ny * q2 : (axis == 1 ? q2 + (long)nx * q1 : q1 + (long)nx * q2);
                const double *t = (s->tab[side] && f < s->tab_n[side]) ? s->tab[side] + 6 * f : NULL;
                if (t && t[5] != 0.0) { cv_t gv = {t[0], t[1], t[2], t[3], t[4]}; s->u[IDX(s, gh[0], gh[1], gh[2])] = gv; }
                else s->u[IDX(s, gh[0], gh[1], gh[2])] = in;
                continue;
            }
            s->u[IDX(s, gh[0], gh[1], gh[2])] = bc_ghost(s, side, &in);
        }
    }
    if (s->c.parabolic_inflow) {                                    /* cylinder main.rs:151-177 */
        double um = s->c.inflow_umax, h = s->c.inflow_h, rho = s->c.inflow_rho;
        for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) {
            long first = IDX(s, NG, j + NG, k + NG);
            double y = centroid(s, 1, j + NG);
            double uy = 4.0 * um * y * (h - y) / (h * h);
            double pin = pressure(&s->u[first], s->c.gamma);
            double t = pin / (rho * 287.058);
            pv_t pv = {rho, uy, 0.0, 0.0, pin, t};
            cv_t c = prim_to_cons(&pv, s->c.gamma);
            for (int gi = 0; gi < NG; gi++) s->u[IDX(s, gi, j + NG, k + NG)] = c;
        }
    }
}

/* ---- IBM (crates/cfd_ibm/src/lib.rs) ---- */
static double cyl_sd(const double *cy, double x, double y) { double dx = x - cy[0], dy = y - cy[1]; return sqrt(dx * dx + dy * dy) - cy[2]; }
static void cyl_normal(const double *cy, double x, double y, double *n) {
    double dx = x - cy[0], dy = y - cy[1];
    double r = sqrt(dx * dx + dy * dy);
    if (r < 1e-12) { n[0] = 1.0; n[1] = 0.0; n[2] = 0.0; } else { n[0] = dx / r; n[1] = dy / r; n[2] = 0.0; }
}
static double sph_sd(const double *b, const double *p) {              /* sphere.rs:43-48 */
    double d0 = p[0] - b[0], d1 = p[1] - b[1], d2 = p[2] - b[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2) - b[3];
}
static void sph_normal(const double *b, const double *p, double *n) {  /* sphere.rs:50-59 */
    double d0 = p[0] - b[0], d1 = p[1] - b[1], d2 = p[2] - b[2];
    double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    if (r < 1e-12) { n[0] = 1.0; n[1] = 0.0; n[2] = 0.0; } else { n[0] = d0 / r; n[1] = d1 / r; n[2] = d2 / r; }
}
/* sphere_markers (sphere.rs:213-237): golden-angle spherical Fibonacci lattice, out[7 m..] =
   position, radial normal, area 4 pi R^2 / n. */
void orc_sphere_markers(const double *b, int n, double *out) {
    const double PI = 3.14159265358979323846;
    double golden = PI * (3.0 - sqrt(5.0));
    double area = 4.0 * PI * b[3] * b[3] / (double)n;
    for (int m = 0; m < n; m++) {
        double zr = 1.0 - 2.0 * ((double)m + 0.5) / (double)n;
        double rr = sqrt(fmax(1.0 - zr * zr, 0.0));
        double phi = golden * (double)m;
        double dir[3] = {rr * cos(phi), rr * sin(phi), zr};
        double *o = out + 7 * m;
        o[0] = b[0] + b[3] * dir[0]; o[1] = b[1] + b[3] * dir[1]; o[2] = b[2] + b[3] * dir[2];
        o[3] = dir[0]; o[4] = dir[1]; o[5] = dir[2]; o[6] = area;
    }
}
double orc_sphere_sd(const double *b, const double *p) { return sph_sd(b, p); }
void orc_sphere_normal(const double *b, const double *p, double *n) { sph_normal(b, p, n); }
static int locate(const orc_sim *s, const double *p, int lo[3], double t[3]) {
    int n[3] = {s->c.nx, s->c.ny, s->c.nz};
    for (int d = 0; d < 3; d++) {
        double q = (p[d] - s->c.lo[d]) / s->h[d] - 0.5;
        double f = floor(q);
        int i0 = (int)f;
        if (i0 < -1 || i0 > n[d] - 1) return 0;
        lo[d] = i0 + NG; t[d] = q - f;
    }
    return 1;
}
static int sample_trilinear(const orc_sim *s, const double *pt, pv_t *out) {  /* cfd_ibm:106-149 */
    int lo[3]; double t[3];
    if (!locate(s, pt, lo, t)) return 0;
    double acc[6] = {0, 0, 0, 0, 0, 0}, tw = 0.0;
    for (int di = 0; di < 2; di++) {
        double wx = di == 0 ? 1.0 - t[0] : t[0];
        for (int dj = 0; dj < 2; dj++) {
            double wy = dj == 0 ? 1.0 - t[1] : t[1];
            for (int dk = 0; dk < 2; dk++) {
                double wz = dk == 0 ? 1.0 - t[2] : t[2];
                long c = IDX(s, lo[0] + di, lo[1] + dj, lo[2] + dk);
                if (s->solid[c]) continue;
                const cv_t *u = &s->u[c];
                if (!isfinite(u->r) || u->r <= 0.0) continue;
                double w = wx * wy * wz;
                pv_t p = cons_to_prim(u, s->c.gamma, s->c.r_gas);
                acc[0] += w * p.r; acc[1] += w * p.u; acc[2] += w * p.v; acc[3] += w * p.w; acc[4] += w * p.p; acc[5] += w * p.t;
                tw += w;
            }
        }
    }
    if (tw < 1e-12) return 0;
    double inv = 1.0 / tw;
    pv_t o = {acc[0] * inv, acc[1] * inv, acc[2] * inv, acc[3] * inv, acc[4] * inv, acc[5] * inv};
    *out = o;
    return 1;
}
static int n_bodies(const orc_sim *s) { return s->c.n_cyl + s->c.n_sph; }
static void ibm_setup(orc_sim *s) {
    if (n_bodies(s) <= 0) return;
    int sph = s->c.n_sph > 0;
    double g = s->c.gamma;
    cv_t fs = {s->c.freestream[0], s->c.freestream[1], s->c.freestream[2], s->c.freestream[3], s->c.freestream[4]};
    cv_t stagnant = {fs.r, 0.0, 0.0, 0.0, fs.re - kinetic(&fs)};   /* cfd_ibm:663-667 */
    (void)g;
    /* refresh_cylinder_mask (cfd_ibm:155-183) */
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++) {
        long c = IDX(s, i + NG, j + NG, k + NG);
        double x = centroid(s, 0, i + NG), y = centroid(s, 1, j + NG);
        double p3[3] = {x, y, centroid(s, 2, k + NG)};
        int inside = 0;
        if (sph) { for (int b = 0; b < s->c.n_sph; b++) if (sph_sd(s->c.sph[b], p3) < 0.0) inside = 1; }   /* sphere.rs:82-111 */
        else for (int b = 0; b < s->c.n_cyl; b++) if (cyl_sd(s->c.cyl[b], x, y) < 0.0) inside = 1;
        if (inside && !s->solid[c]) { s->solid[c] = 1; s->u[c] = stagnant; memset(&s->rhs[c], 0, sizeof(cv_t)); }
        else if (!inside && s->solid[c]) { s->solid[c] = 0; s->u[c] = fs; memset(&s->rhs[c], 0, sizeof(cv_t)); }
    }
    if (s->ibm_ready) return;   /* static bodies: the ghost list and markers do not change */
    /* compute_cylinder_ghosts (cfd_ibm:187-254) */
    free(s->ghost_idx); free(s->ghost_img);
    s->ghost_idx = (long *)malloc(sizeof(long) * s->ntot);
    s->ghost_img = malloc(sizeof(*s->ghost_img) * s->ntot);
    s->n_ghost = 0;
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++) {
        int ir = i + NG, jr = j + NG, kr = k + NG;
        long c = IDX(s, ir, jr, kr);
        if (!s->solid[c]) continue;
        long nb[6] = {IDX(s, ir + 1, jr, kr), IDX(s, ir - 1, jr, kr), IDX(s, ir, jr + 1, kr), IDX(s, ir, jr - 1, kr), IDX(s, ir, jr, kr + 1), IDX(s, ir, jr, kr - 1)};
        int anyfluid = 0;
        for (int q = 0; q < 6; q++) if (!s->solid[nb[q]]) anyfluid = 1;
        if (!anyfluid) continue;
        double cc[3] = {centroid(s, 0, ir), centroid(s, 1, jr), centroid(s, 2, kr)};
        double best_abs = 1.79769313486231570815e308, best_signed = 0.0, bn[3] = {1.0, 0.0, 0.0};
        if (sph) {                                                     /* compute_sphere_ghosts, sphere.rs:116-187 */
            for (int b = 0; b < s->c.n_sph; b++) {
                double sd = sph_sd(s->c.sph[b], cc);
                if (fabs(sd) < best_abs) { best_abs = fabs(sd); best_signed = sd; sph_normal(s->c.sph[b], cc, bn); }
            }
        } else
        for (int b = 0; b < s->c.n_cyl; b++) {
            double sd = cyl_sd(s->c.cyl[b], cc[0], cc[1]);
            if (fabs(sd) < best_abs) { best_abs = fabs(sd); best_signed = sd; cyl_normal(s->c.cyl[b], cc[0], cc[1], bn); }
        }
        double push = 2.0 * (-best_signed);
        s->ghost_idx[s->n_ghost] = c;
        s->ghost_img[s->n_ghost][0] = cc[0] + push * bn[0];
        s->ghost_img[s->n_ghost][1] = cc[1] + push * bn[1];
        s->ghost_img[s->n_ghost][2] = cc[2] + push * bn[2];
        s->n_ghost++;
    }
    if (sph) {                                                        /* compute_sphere_markers, sphere.rs:189-207 */
        int n = s->c.sph_markers > 12 ? s->c.sph_markers : 12;
        free(s->mark);
        s->mark = malloc(sizeof(*s->mark) * (size_t)s->c.n_sph * n);
        s->n_mark = 0;
        for (int b = 0; b < s->c.n_sph; b++) { orc_sphere_markers(s->c.sph[b], n, &s->mark[s->n_mark][0]); s->n_mark += n; }
        s->ibm_ready = 1;
        return;
    }
    /* compute_cylinder_markers (cfd_ibm:566-598) */
    int nk = s->c.nz;
    double z_lo = centroid(s, 2, NG) - 0.5 * s->h[2];
    double z_hi = centroid(s, 2, NG + nk - 1) + 0.5 * s->h[2];
    double lz = z_hi - z_lo;
    int n_theta = 64, n_z = nk > 1 ? nk : 1;
    const double PI = 3.14159265358979323846;
    free(s->mark);
    s->mark = malloc(sizeof(*s->mark) * (size_t)s->c.n_cyl * n_theta * n_z);
    s->n_mark = 0;
    for (int b = 0; b < s->c.n_cyl; b++) {
        double area = 2.0 * PI * s->c.cyl[b][2] * lz / (double)(n_theta * n_z);
        for (int it = 0; it < n_theta; it++) {
            double th = 2.0 * PI * ((double)it + 0.5) / (double)n_theta;
            double nx = cos(th), ny = sin(th);
            for (int iz = 0; iz < n_z; iz++) {
                double z = z_lo + ((double)iz + 0.5) * lz / (double)n_z;
                double *m = s->mark[s->n_mark++];
                m[0] = s->c.cyl[b][0] + s->c.cyl[b][2] * nx; m[1] = s->c.cyl[b][1] + s->c.cyl[b][2] * ny; m[2] = z;
                m[3] = nx; m[4] = ny; m[5] = 0.0; m[6] = area;
            }
        }
    }
    s->ibm_ready = 1;
}
long orc_ibm_ghosts(const orc_sim *s, long *idx, double *img) {
    if (idx) for (long q = 0; q < s->n_ghost; q++) { idx[q] = s->ghost_idx[q]; img[3 * q] = s->ghost_img[q][0]; img[3 * q + 1] = s->ghost_img[q][1]; img[3 * q + 2] = s->ghost_img[q][2]; }
    return s->n_ghost;
}
void orc_get_solid(const orc_sim *s, unsigned char *dst) { memcpy(dst, s->solid, s->ntot); }
static void apply_ibm_ghost_bc(orc_sim *s) {                         /* cfd_ibm:258-289 */
    for (long q = 0; q < s->n_ghost; q++) {
        pv_t im;
        if (!sample_trilinear(s, s->ghost_img[q], &im)) continue;
        pv_t m = {im.r, 2.0 * 0.0 - im.u, 2.0 * 0.0 - im.v, 2.0 * 0.0 - im.w, im.p, im.t};
        s->u[s->ghost_idx[q]] = prim_to_cons(&m, s->c.gamma);
    }
}
static void marker_force_system(orc_sim *s) {                       /* cfd_ibm:442-480,601-615 */
    double fp[3] = {0, 0, 0}, fv[3] = {0, 0, 0};
    double off = s->h[0];
    for (long q = 0; q < s->n_mark; q++) {
        const double *m = s->mark[q];
        double sp[3] = {m[0] + off * m[3], m[1] + off * m[4], m[2] + off * m[5]};
        pv_t im;
        if (!sample_trilinear(s, sp, &im)) continue;
        if (!(isfinite(im.p) && isfinite(im.t) && isfinite(im.u) && isfinite(im.v) && isfinite(im.w))) continue;
        double us[3] = {im.u - 0.0, im.v - 0.0, im.w - 0.0};
        double udn = us[0] * m[3] + us[1] * m[4] + us[2] * m[5];
        double mu = visc_mu(&s->c, im.t);
        for (int d = 0; d < 3; d++) {
            double ut = us[d] - udn * m[3 + d];
            fp[d] += -im.p * m[3 + d] * m[6];
            fv[d] += mu * ut / off * m[6];
        }
    }
    for (int d = 0; d < 3; d++) { s->marker_force[d] = fp[d]; s->marker_force[3 + d] = fv[d]; }
}
static void wall_force_system(orc_sim *s) {                         /* cfd_ibm:309-362 */
    double fp[3] = {0, 0, 0}, fv[3] = {0, 0, 0};
    double g = s->c.gamma;
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++) {
        int ir = i + NG, jr = j + NG, kr = k + NG;
        long c = IDX(s, ir, jr, kr);
        if (!s->solid[c]) continue;
        double dx = spacing(s, 0, i), dy = spacing(s, 1, j), dz = spacing(s, 2, k);
        double fa[3] = {dy * dz, dx * dz, dx * dy}, dn[3] = {dx, dy, dz};
        long nb[6] = {IDX(s, ir + 1, jr, kr), IDX(s, ir - 1, jr, kr), IDX(s, ir, jr + 1, kr), IDX(s, ir, jr - 1, kr), IDX(s, ir, jr, kr + 1), IDX(s, ir, jr, kr - 1)};
        int dims[6] = {0, 0, 1, 1, 2, 2};
        double sg[6] = {1.0, -1.0, 1.0, -1.0, 1.0, -1.0};
        for (int q = 0; q < 6; q++) {
            if (s->solid[nb[q]]) continue;
            const cv_t *u = &s->u[nb[q]];
            int dim = dims[q];
            fp[dim] -= pressure(u, g) * sg[q] * fa[dim];
            pv_t p = cons_to_prim(u, g, s->c.r_gas);
            double mu = visc_mu(&s->c, p.t);
            double vel[3] = {p.u, p.v, p.w};
            for (int td = 0; td < 3; td++) if (td != dim) fv[td] += mu * vel[td] / (0.5 * dn[dim]) * fa[dim];
        }
    }
    for (int d = 0; d < 3; d++) { s->wall_force[d] = fp[d]; s->wall_force[3 + d] = fv[d]; }
}
void orc_marker_force(orc_sim *s, double *out) { marker_force_system(s); memcpy(out, s->marker_force, sizeof(double) * 6); }
void orc_wall_force(orc_sim *s, double *out) { wall_force_system(s); memcpy(out, s->wall_force, sizeof(double) * 6); }

static void zero_rhs(orc_sim *s) { memset(s->rhs, 0, sizeof(cv_t) * s->ntot); }

static void slopes_system(orc_sim *s) {                             /* cfd_solver:233-268 */
    int lim = s->c.limiter;
    memset(s->slope, 0, sizeof(*s->slope) * s->ntot);
#pragma omp parallel for collapse(2) OMP_THREADS(s)
    for (int k = NG; k < NG + s->c.nz; k++) for (int j = NG; j < NG + s->c.ny; j++) for (int i = NG; i < NG + s->c.nx; i++) {
        long c = IDX(s, i, j, k);
        cv_t us = s->u[c];
        long st[3] = {1, s->NX, (long)s->NX * s->NY};
        for (int ax = 0; ax < 3; ax++) {
            cv_t up = s->u[c + st[ax]], un = s->u[c - st[ax]];
            /* neighbor_mean of one cell: ZERO + u, then * (1.0/1) */
            s->slope[c][ax] = limited(cv_sub(up, us), cv_sub(us, un), lim);
        }
    }
}
/* Periodic image of a raw index along every periodic axis (conservative seam mode). */
static long wrap_raw(const orc_sim *s, int i, int j, int k) {
    int n[3] = {s->c.nx, s->c.ny, s->c.nz}; int id[3] = {i, j, k};
    for (int d = 0; d < 3; d++) if (axis_periodic(s, d)) {
        if (id[d] < NG) id[d] += n[d]; else if (id[d] >= NG + n[d]) id[d] -= n[d];
    }
    return IDX(s, id[0], id[1], id[2]);
}
static inline cv_t recon(cv_t u, const cv_t sl[3], double frac_axis, int axis) {  /* cfd_solver:273-286 */
    for (int a = 0; a < 3; a++) {
        double fr = a == axis ? frac_axis : 0.0;
        u = cv_add(u, cv_mul(sl[a], fr));
    }
    return u;
}

/* The flux through one face of the sweep: cell (i, j, k), axis ax, side 0 (+ face) or 1 (- face
 * at the low edge).  Sets *scatter when the serial sweep also adds it to the + neighbour. */
static cv_t sweep_face(const orc_sim *s, int i, int j, int k, int ax, int side, int *scatter) {
    int muscl = s->c.recon == RECON_MUSCL;
    long st[3] = {1, s->NX, (long)s->NX * s->NY};
    int n[3] = {s->c.nx, s->c.ny, s->c.nz};
    int id[3] = {i, j, k};
    long own = IDX(s, i + NG, j + NG, k + NG);
    int sign = side == 0 ? 1 : -1;
    long oth = own + sign * st[ax];
    int other_local = side == 0 ? (id[ax] + 1 < n[ax]) : 0;
    int edge = side == 0 ? (id[ax] + 1 == n[ax]) : 1;
    int patch = edge && !axis_periodic(s, ax);
    cv_t ul = s->u[own], ur = s->u[oth];
    if (muscl && !patch) {
        ul = recon(s->u[own], s->slope[own], 0.5 * sign, ax);
        if (other_local) ur = recon(s->u[oth], s->slope[oth], -0.5 * sign, ax);
        else if (s->c.seam_mode == 1) {
            int o3[3] = {i + NG, j + NG, k + NG}; o3[ax] += sign;
            long w = wrap_raw(s, o3[0], o3[1], o3[2]);
            ur = recon(s->u[w], s->slope[w], -0.5 * sign, ax);
        }
    }
    *scatter = !patch && other_local;
    return flux_normal_axis(s->c.flux, &ul, &ur, s->c.gamma, ax, sign);
}

/* Threaded sweep, bitwise the serial one.  Phase 1 evaluates every face flux; phase 2 forms each
 * cell's right-hand side in the serial sweep's order: the scatters from the - neighbours along z,
 * y, x (visited earlier, in that order), then the cell's own faces x+, x-(low edge), y+, y-, z+, z-. */
static void flux_sweep_threaded(orc_sim *s) {
    int n[3] = {s->c.nx, s->c.ny, s->c.nz};
    long nin = (long)n[0] * n[1] * n[2];
    long edge_n[3] = {(long)n[1] * n[2], (long)n[0] * n[2], (long)n[0] * n[1]};
    for (int d = 0; d < 3; d++) {
        if (!s->fplus[d]) s->fplus[d] = (cv_t *)malloc(sizeof(cv_t) * nin);
        if (!s->fminus[d]) s->fminus[d] = (cv_t *)malloc(sizeof(cv_t) * edge_n[d]);
    }
#pragma omp parallel for collapse(2) OMP_THREADS(s)
    for (int k = 0; k < n[2]; k++) for (int j = 0; j < n[1]; j++) for (int i = 0; i < n[0]; i++) {
        long q = i + (long)n[0] * (j + (long)n[1] * k);
        int id[3] = {i, j, k}, sc;
        for (int ax = 0; ax < 3; ax++) {
            s->fplus[ax][q] = sweep_face(s, i, j, k, ax, 0, &sc);
            if (id[ax] == 0) {
                long e = ax == 0 ? j + (long)n[1] * k : (ax == 1 ? i + (long)n[0] * k : i + (long)n[0] * j);
                s->fminus[ax][e] = sweep_face(s, i, j, k, ax, 1, &sc);
            }
        }
    }
#pragma omp parallel for collapse(2) OMP_THREADS(s)
    for (int k = 0; k < n[2]; k++) for (int j = 0; j < n[1]; j++) for (int i = 0; i < n[0]; i++) {
        long q = i + (long)n[0] * (j + (long)n[1] * k);
        long sq[3] = {1, n[0], (long)n[0] * n[1]};
        int id[3] = {i, j, k};
        /* area / volume of this cell: faces shared along an axis have the same area on both sides */
        const double dx = spacing(s, 0, i), dy = spacing(s, 1, j), dz = spacing(s, 2, k), vol = dx * dy * dz;
        const double area[3] = {dy * dz, dx * dz, dx * dy};
        cv_t acc = s->rhs[IDX(s, i + NG, j + NG, k + NG)];
        for (int ax = 2; ax >= 0; ax--)               /* + faces of the z, y, x neighbours */
            if (id[ax] >= 1) acc = cv_add(acc, cv_mul(s->fplus[ax][q - sq[ax]], area[ax] / vol));
        for (int ax = 0; ax < 3; ax++) {
            double inv = area[ax] / vol;
            acc = cv_sub(acc, cv_mul(s->fplus[ax][q], inv));
            if (id[ax] == 0) {
                long e = ax == 0 ? j + (long)n[1] * k : (ax == 1 ? i + (long)n[0] * k : i + (long)n[0] * j);
                acc = cv_sub(acc, cv_mul(s->fminus[ax][e], inv));
            }
        }
        s->rhs[IDX(s, i + NG, j + NG, k + NG)] = acc;
    }
}

static void flux_sweep(orc_sim *s) {                                /* cfd_solver:93-122 / 293-339 */
    if (s->threads > 1) { flux_sweep_threaded(s); return; }
    int muscl = s->c.recon == RECON_MUSCL;
    int scheme = s->c.flux;
    double g = s->c.gamma;
    long st[3] = {1, s->NX, (long)s->NX * s->NY};
    int n[3] = {s->c.nx, s->c.ny, s->c.nz};
    for (int k = 0; k < n[2]; k++) for (int j = 0; j < n[1]; j++) for (int i = 0; i < n[0]; i++) {
        int id[3] = {i, j, k};
        long own = IDX(s, i + NG, j + NG, k + NG);
        /* cfd_solver:418-424: the owner's spacings give the face area and its volume */
        const double dx = spacing(s, 0, i), dy = spacing(s, 1, j), dz = spacing(s, 2, k), vol = dx * dy * dz;
        const double area[3] = {dy * dz, dx * dz, dx * dy};
        for (int ax = 0; ax < 3; ax++) {
            double inv = area[ax] / vol;
            for (int side = 0; side < 2; side++) {        /* 0: + face; 1: - face at the low edge */
                if (side == 1 && id[ax] != 0) continue;
                int sign = side == 0 ? 1 : -1;
                long oth = own + sign * st[ax];
                int other_local = side == 0 ? (id[ax] + 1 < n[ax]) : 0;
                int edge = side == 0 ? (id[ax] + 1 == n[ax]) : 1;
                int patch = edge && !axis_periodic(s, ax);  /* physical boundary face */
                cv_t ul = s->u[own], ur = s->u[oth];
                if (muscl && !patch) {
                    ul = recon(s->u[own], s->slope[own], 0.5 * sign, ax);
                    if (other_local) ur = recon(s->u[oth], s->slope[oth], -0.5 * sign, ax);
                    else if (s->c.seam_mode == 1) {
                        int o3[3] = {i + NG, j + NG, k + NG}; o3[ax] += sign;
                        long w = wrap_raw(s, o3[0], o3[1], o3[2]);
                        ur = recon(s->u[w], s->slope[w], -0.5 * sign, ax);
                    }
                }
                cv_t fn = flux_normal_axis(scheme, &ul, &ur, g, ax, sign);
                s->rhs[own] = cv_sub(s->rhs[own], cv_mul(fn, inv));
                if (!patch && other_local) {
                    /* the face's area over the neighbour's volume (cfd_solver:432-434) */
                    int jo = j + (ax == 1), ko = k + (ax == 2);
                    double vol_o = spacing(s, 0, i + (ax == 0)) * spacing(s, 1, jo) * spacing(s, 2, ko);
                    s->rhs[oth] = cv_add(s->rhs[oth], cv_mul(fn, area[ax] / vol_o));
                }
            }
        }
    }
}

static inline double d2_pre(double um, double u0, double up, double ihm, double ihp, double ihs) {
    return ((up - u0) * ihp - (u0 - um) * ihm) * ihs;               /* cfd_solver:449-453 */
}
static void viscous_system(orc_sim *s) {                            /* cfd_solver:463-560 */
    double g = s->c.gamma, R = s->c.r_gas;
#pragma omp parallel for OMP_THREADS(s)
    for (long c = 0; c < s->ntot; c++) s->prim[c] = cons_to_prim(&s->u[c], g, R);
    int nx = s->c.nx, ny = s->c.ny, nz = s->c.nz;
    double dx = s->h[0];
    double dxi2 = 1.0 / (dx * dx);
    double *yim = malloc(sizeof(double) * ny * 3), *yip = yim + ny, *yis = yim + 2 * ny;
    double *zim = malloc(sizeof(double) * nz * 3), *zip = zim + nz, *zis = zim + 2 * nz;
    for (int j = 0; j < ny; j++) {
        int jr = j + NG;
        double hm = centroid(s, 1, jr) - centroid(s, 1, jr - 1), hp = centroid(s, 1, jr + 1) - centroid(s, 1, jr);
        yim[j] = 1.0 / hm; yip[j] = 1.0 / hp; yis[j] = 2.0 / (hm + hp);
    }
    for (int k = 0; k < nz; k++) {
        int kr = k + NG;
        double hm = centroid(s, 2, kr) - centroid(s, 2, kr - 1), hp = centroid(s, 2, kr + 1) - centroid(s, 2, kr);
        zim[k] = 1.0 / hm; zip[k] = 1.0 / hp; zis[k] = 2.0 / (hm + hp);
    }
#pragma omp parallel for collapse(2) OMP_THREADS(s)
    for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) {
        int ir = i + NG, jr = j + NG, kr = k + NG;
        long c = IDX(s, ir, jr, kr);
        if (s->solid[c]) continue;
        pv_t ps = s->prim[c];
        double mu = visc_mu(&s->c, ps.t);
        double kap = conductivity(mu, g, R);
        long nb[6] = {IDX(s, ir + 1, jr, kr), IDX(s, ir - 1, jr, kr), IDX(s, ir, jr + 1, kr), IDX(s, ir, jr - 1, kr), IDX(s, ir, jr, kr + 1), IDX(s, ir, jr, kr - 1)};
        double q[6][4];
        for (int t = 0; t < 6; t++) {
            if (s->solid[nb[t]]) { q[t][0] = 0.0; q[t][1] = 0.0; q[t][2] = 0.0; q[t][3] = ps.t; }
            else { const pv_t *p = &s->prim[nb[t]]; q[t][0] = p->u; q[t][1] = p->v; q[t][2] = p->w; q[t][3] = p->t; }
        }
        double vel[4] = {ps.u, ps.v, ps.w, ps.t};
        double lap[4];
        for (int m = 0; m < 4; m++) {
            double d2x = (q[0][m] - 2.0 * vel[m] + q[1][m]) * dxi2;
            double d2y = d2_pre(q[3][m], vel[m], q[2][m], yim[j], yip[j], yis[j]);
            double d2z = d2_pre(q[5][m], vel[m], q[4][m], zim[k], zip[k], zis[k]);
            lap[m] = d2x + d2y + d2z;
        }
        double fx = mu * lap[0], fy = mu * lap[1], fz = mu * lap[2];
        double en = vel[0] * fx + vel[1] * fy + vel[2] * fz + kap * lap[3];
        cv_t add = {0.0, fx, fy, fz, en};
        s->rhs[c] = cv_add(s->rhs[c], add);
    }
    free(yim); free(zim);
}
static void body_force_system(orc_sim *s) {  /* poiseuille_channel_validation.rs:105-120 */
    double gx = s->c.body_gx;
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++) {
        long c = IDX(s, i + NG, j + NG, k + NG);
        double rho = s->u[c].r, ru = s->u[c].ru;
        s->rhs[c].ru += rho * gx;
        s->rhs[c].re += gx * ru;
    }
}
static void zero_solid_rhs(orc_sim *s) {                             /* cfd_ibm:293-301 */
    for (long c = 0; c < s->ntot; c++) if (s->solid[c]) memset(&s->rhs[c], 0, sizeof(cv_t));
}
static int update_system(orc_sim *s) {                              /* integrate.rs:46-101; cfd_solver:607-621 */
    double dt = s->dt;
    int completed = 0;
    if (s->c.integrator == INTEG_EULER) {
#pragma omp parallel for OMP_THREADS(s)
        for (long c = 0; c < s->ntot; c++) s->u[c] = cv_add(s->u[c], cv_mul(s->rhs[c], dt));
        s->stage = STAGE_FIRST; completed = 1;
    } else if (s->stage == STAGE_FIRST) {
        memcpy(s->u0, s->u, sizeof(cv_t) * s->ntot);
#pragma omp parallel for OMP_THREADS(s)
        for (long c = 0; c < s->ntot; c++) s->u[c] = cv_add(s->u[c], cv_mul(s->rhs[c], dt));
        s->stage = STAGE_SECOND;
    } else if (s->stage == STAGE_SECOND) {
#pragma omp parallel for OMP_THREADS(s)
        for (long c = 0; c < s->ntot; c++)
            s->u[c] = cv_add(cv_mul(s->u0[c], 0.75), cv_mul(cv_add(s->u[c], cv_mul(s->rhs[c], dt)), 0.25));
        s->stage = STAGE_THIRD;
    } else {
#pragma omp parallel for OMP_THREADS(s)
        for (long c = 0; c < s->ntot; c++)
            s->u[c] = cv_add(cv_mul(s->u0[c], 1.0 / 3.0), cv_mul(cv_add(s->u[c], cv_mul(s->rhs[c], dt)), 2.0 / 3.0));
        s->stage = STAGE_FIRST; completed = 1;
    }
    if (completed) { s->step += 1; s->time += s->dt; }
    return completed;
}
static void sponge_system(orc_sim *s) {                             /* cylinder main.rs:186-229 */
    if (s->stage != STAGE_FIRST) return;
    double g = s->c.gamma;
    double dt = s->dt, x0 = s->c.sp_x0, x1 = s->c.sp_x1, sig = s->c.sp_sigma;
    double um = s->c.sp_umax, h = s->c.sp_h, rho = s->c.sp_rho, p = s->c.sp_p;
    double inv_w = 1.0 / (x1 - x0);
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++) {
        long idx = IDX(s, i + NG, j + NG, k + NG);
        double cx = centroid(s, 0, i + NG), cy = centroid(s, 1, j + NG);
        if (cx < x0) continue;
        double sc = (cx - x0) * inv_w; sc = sc < 0.0 ? 0.0 : (sc > 1.0 ? 1.0 : sc);
        double f = fmin(sig * sc * sc * dt, 1.0);
        cv_t tg;
        if (s->c.sp_uniform) {                                       /* sphere main.rs:119-160 */
            cv_t t = {s->c.sp_target[0], s->c.sp_target[1], s->c.sp_target[2], s->c.sp_target[3], s->c.sp_target[4]};
            tg = t;
        } else {
            double uy = 4.0 * um * cy * (h - cy) / (h * h);
            pv_t pv = {rho, uy, 0.0, 0.0, p, p / (rho * 287.058)};
            tg = prim_to_cons(&pv, g);
        }
        cv_t *u = &s->u[idx];
        u->r += f * (tg.r - u->r); u->ru += f * (tg.ru - u->ru); u->rv += f * (tg.rv - u->rv);
        u->rw += f * (tg.rw - u->rw); u->re += f * (tg.re - u->re);
    }
}

/* One app.run(): Setup -> Halo -> Boundary -> PreFlux -> Flux -> PostFlux -> Update -> PostUpdate -> Output. */
int orc_run(orc_sim *s) {
    cfl_system(s);                       /* Setup: SolverPlugin's CFL (registered first) */
    ibm_setup(s);                        /* Setup: CylinderIbmPlugin mask / ghosts / markers */
    halo_system(s);                      /* Halo */
    boundary_system(s);                  /* Boundary (+ parabolic inflow after the stock fill) */
    zero_rhs(s);                         /* PreFlux */
    if (n_bodies(s) > 0) apply_ibm_ghost_bc(s);
    if (s->c.recon == RECON_MUSCL) slopes_system(s);
    flux_sweep(s);                       /* Flux: inviscid, then viscous */
    if (s->c.viscous) viscous_system(s);
    if (s->c.body_force) body_force_system(s);   /* PostFlux */
    if (n_bodies(s) > 0) zero_solid_rhs(s);
    int done = update_system(s);         /* Update */
    if (s->c.sponge) sponge_system(s);   /* PostUpdate */
    if (n_bodies(s) > 0) marker_force_system(s);  /* Output */
    return done;
}
/* One full time step (3 substages for RK3, 1 for Euler). */
void orc_step(orc_sim *s) { int n = s->c.integrator == INTEG_RK3 ? 3 : 1; for (int q = 0; q < n; q++) orc_run(s); }
/* The harness loop `loop { if t >= t_final break; step }` (examples/isentropic_vortex/main.rs:179-187). */
long orc_advance_to(orc_sim *s, double t_final, long max_steps) {
    long k = 0;
    while (s->time < t_final && k < max_steps) { orc_step(s); k++; }
    return k;
}
void orc_advance_steps(orc_sim *s, long n) { for (long k = 0; k < n; k++) orc_step(s); }
/* Right-hand side of the current state (after one substage's ghost fill), for unit checks. */
void orc_get_rhs_interior(const orc_sim *s, double *dst) {
    long q = 0;
    for (int k = 0; k < s->c.nz; k++) for (int j = 0; j < s->c.ny; j++) for (int i = 0; i < s->c.nx; i++, q++)
        memcpy(dst + 5 * q, &s->rhs[IDX(s, i + NG, j + NG, k + NG)], sizeof(cv_t));
}
