// efvm_stage3d.hip -- 3-D structured meshes (UniformMesh with nz >= 1): the FP64
// parity path and its FP32 twin, compiled with -ffp-contract=off.
//
// Device layout: the reference's raw padded array (ng = 2 ghost layers on every
// axis, crates/cfd_state/src/lib.rs:152-159), five SoA planes (rho, rho_u, rho_v,
// rho_w, rho_E), x fastest, so ghost layers, halo copies and the immersed-boundary
// mirror address cells exactly as the reference's systems do.
//
// One substage (one app.run()):
//   [first substage of a step] CFL scan -> per-block maxima -> lam (cfd_solver:626-677)
//   Halo:     periodic copy, one launch per periodic axis, x then y then z
//   Boundary: first ghost layer of every physical side from its owner (cfd_boundary:574-601)
//   PreFlux:  IBM image-point mirror into the solid ghost cells (cfd_ibm:258-289)
//   Flux .. Update: one fused kernel over every raw cell.  An interior cell evaluates
//     the fluxes of its faces in the order the reference's face sweep scatters them
//     (cfd_solver:93-122, 293-339): the + faces of its lower z, y, x neighbours, then
//     its own + face (and - face on a low boundary) per axis; each face is evaluated
//     by the two cells it separates with the same operands, so both see the same
//     value.  Then viscous (cfd_solver:463-560), body force, solid-rhs zero, and the
//     RK substep (integrate.rs:69-101), which the reference applies to every raw
//     cell, ghosts included with rhs = 0.
//   PostUpdate: sponge after a completed step.
#include "efvm_stage3d.h"

namespace efvm {

template <class T> using V5 = C5<T>;

template <class T>
__device__ __forceinline__ V5<T> ld5(const T* const* U, long q) {
    return {U[0][q], U[1][q], U[2][q], U[3][q], U[4][q]};
}
template <class T>
__device__ __forceinline__ void st5(T* const* U, long q, const V5<T>& v) {
    U[0][q] = v.r; U[1][q] = v.ru; U[2][q] = v.rv; U[3][q] = v.rw; U[4][q] = v.re;
}
template <class T>
__device__ __forceinline__ V5<double> to_d(const V5<T>& v) {
    return {double(v.r), double(v.ru), double(v.rv), double(v.rw), double(v.re)};
}
template <class T>
__device__ __forceinline__ V5<T> from_d(const V5<double>& v) {
    return {T(v.r), T(v.ru), T(v.rv), T(v.rw), T(v.re)};
}

__device__ __forceinline__ double dt_of(const Args3& a) {
    if (a.fixed_dt > 0.0) return a.fixed_dt;
    return a.cfl / a.clock->lam[a.parity];
}

// ---------------------------------------------------------------------------
// Halo: ghost slabs of periodic axis d take the interior cell n[d] away along d
// (the other two indices as they are, ghosts of earlier axes included).  An axis of
// one cell has its outer ghost layer two cells away from any interior cell: it takes
// the periodic image too (the reference copies a not yet refreshed ghost there; no
// stencil, sample or sum reads that layer).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void halo3_kernel(Args3 a, int d) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nraw) return;
    const int i = (int)(q % a.NX);
    const long jk = q / a.NX;
    const int j = (int)(jk % a.NY), k = (int)(jk / a.NY);
    const int id[3] = {i, j, k};
    const int n[3] = {a.nx, a.ny, a.nz};
    const int r = id[d];
    if (r >= NGH && r < NGH + n[d]) return;
    const long st[3] = {1, a.NX, (long)a.NX * a.NY};
    int rs = (r - NGH) % n[d];
    if (rs < 0) rs += n[d];
    const long src = q + (long)(rs + NGH - r) * st[d];
    T* const* U = (T* const*)a.uin;
    st5(U, q, ld5((const T* const*)a.uin, src));
}

// Ghost state of one physical side from its owner cell (cfd_boundary:173-362).
__device__ __forceinline__ V5<double> bc_ghost5(const Args3& a, int side, const V5<double>& in) {
    const double g = a.gamma;
    const double* par = a.bc_par[side];
    const int axis = side / 2;
    V5<double> o = in;
    switch (a.bc_kind[side]) {
    case BC_SLIP:
    case BC_SYMMETRY:
        if (axis == 0) o.ru = -o.ru; else if (axis == 1) o.rv = -o.rv; else o.rw = -o.rw;
        return o;
    case BC_NOSLIP:
        o.ru = -o.ru; o.rv = -o.rv; o.rw = -o.rw;
        return o;
    case BC_SUP_INFLOW:
        return {a.bc_cons[side][0], a.bc_cons[side][1], a.bc_cons[side][2], a.bc_cons[side][3], a.bc_cons[side][4]};
    case BC_SUB_INFLOW: {
        const double pin = ref::pressure(in, g);
        P6<double> p = {par[0], par[1], par[2], par[3], pin, par[4]};
        return ref::prim_to_cons(p, g);
    }
    case BC_PRESSURE_OUTLET: {
        P6<double> p = ref::cons_to_prim(in, g, a.r_gas);
        p.p = par[0];
        p.t = par[0] / (p.r * 287.058);
        return ref::prim_to_cons(p, g);
    }
    default:
        return o;
    }
}

// Boundary: blockIdx.y = side; one thread per face of that side.  The six sides
// write disjoint ghost cells and read interior cells only.
template <class T>
__global__ __launch_bounds__(256) void bc3_kernel(Args3 a) {
    const int side = blockIdx.y;
    if (a.bc_kind[side] == BC_PERIODIC) return;
    const int axis = side / 2, hi = side & 1;
    const int n[3] = {a.nx, a.ny, a.nz};
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    const long nf = (long)n[a1] * n[a2];
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    int own[3], gh[3];
    own[a1] = (int)(f % n[a1]) + NGH;
    own[a2] = (int)(f / n[a1]) + NGH;
    gh[a1] = own[a1];
    gh[a2] = own[a2];
    own[axis] = hi ? NGH + n[axis] - 1 : NGH;
    gh[axis] = hi ? NGH + n[axis] : NGH - 1;
    const long qo = own[0] + (long)a.NX * (own[1] + (long)a.NY * own[2]);
    const long qg = gh[0] + (long)a.NX * (gh[1] + (long)a.NY * gh[2]);
    const V5<double> in = to_d(ld5((const T* const*)a.uin, qo));
    if (a.bc_kind[side] == BC_TABLE) {
        // FromOutside (crates/cfd_boundary/src/lib.rs:135-144): the supplied state where one
        // is given, else the interior copy.  Table index over the side, the lower-numbered
        // in-plane axis fastest: x side j + ny k, y side i + nx k, z side i + nx j.
        const int i = own[0] - NGH, j = own[1] - NGH, k = own[2] - NGH;
        const long t = axis == 0 ? j + (long)a.ny * k : (axis == 1 ? i + (long)a.nx * k : i + (long)a.nx * j);
        const double* e = a.tab[side] ? a.tab[side] + 6 * t : nullptr;
        const V5<double> o = (e && e[5] != 0.0) ? V5<double>{e[0], e[1], e[2], e[3], e[4]} : in;
        st5((T* const*)a.uin, qg, from_d<T>(o));
        return;
    }
    st5((T* const*)a.uin, qg, from_d<T>(bc_ghost5(a, side, in)));
}

// ---------------------------------------------------------------------------
// Immersed boundary (crates/cfd_ibm/src/lib.rs:106-149, 258-289): trilinear
// primitive sample at a point, skipping solid, non-finite and rho <= 0 corners and
// renormalising by the weight kept.  locate(): s = (p - lo)/h - 0.5, i0 = floor(s).
// ---------------------------------------------------------------------------
template <class T>
__device__ bool sample3(const Args3& a, const T* const* U, const double* pt, P6<double>& out) {
    const int n[3] = {a.nx, a.ny, a.nz};
    int lo[3];
    double t[3];
    for (int d = 0; d < 3; d++) {
        const double s = (pt[d] - a.lo[d]) / a.h[d] - 0.5;
        if (!(s > -1.0e9 && s < 1.0e9)) return false;
        const double f = floor(s);
        const int i0 = (int)f;
        if (i0 < -1 || i0 > n[d] - 1) return false;
        lo[d] = i0 + NGH;
        t[d] = s - f;
    }
    double acc[6] = {0, 0, 0, 0, 0, 0}, tw = 0.0;
    for (int di = 0; di < 2; di++) {
        const double wx = di == 0 ? 1.0 - t[0] : t[0];
        for (int dj = 0; dj < 2; dj++) {
            const double wy = dj == 0 ? 1.0 - t[1] : t[1];
            for (int dk = 0; dk < 2; dk++) {
                const double wz = dk == 0 ? 1.0 - t[2] : t[2];
                const long c = (lo[0] + di) + (long)a.NX * ((lo[1] + dj) + (long)a.NY * (lo[2] + dk));
                if (a.solid && a.solid[c]) continue;
                const V5<double> u = to_d(ld5(U, c));
                if (!isfinite(u.r) || u.r <= 0.0) continue;
                const double w = wx * wy * wz;
                const P6<double> p = ref::cons_to_prim(u, a.gamma, a.r_gas);
                acc[0] += w * p.r; acc[1] += w * p.u; acc[2] += w * p.v; acc[3] += w * p.w; acc[4] += w * p.p; acc[5] += w * p.t;
                tw += w;
            }
        }
    }
    if (tw < 1e-12) return false;
    const double inv = 1.0 / tw;
    out = {acc[0] * inv, acc[1] * inv, acc[2] * inv, acc[3] * inv, acc[4] * inv, acc[5] * inv};
    return true;
}

// Mirror of each solid ghost cell: (rho, 2 u_wall - u, p) at its image point, u_wall = 0
// for a fixed body.  Samples skip solid cells and only solid cells are written, so the
// ghosts do not depend on one another.
template <class T>
__global__ void mirror3_kernel(Args3 a, int n, const long* __restrict__ idx, const double* __restrict__ img) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    P6<double> im;
    if (!sample3<T>(a, (const T* const*)a.uin, img + 3 * q, im)) return;
    const P6<double> m = {im.r, 2.0 * 0.0 - im.u, 2.0 * 0.0 - im.v, 2.0 * 0.0 - im.w, im.p, im.t};
    st5((T* const*)a.uin, idx[q], from_d<T>(ref::prim_to_cons(m, a.gamma)));
}

// Cells inside a body take the stagnant freestream state (refresh_*_mask, first Setup).
template <class T>
__global__ __launch_bounds__(256) void flip3_kernel(Args3 a, const double* __restrict__ stagnant) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nraw || !a.solid[q]) return;
    st5((T* const*)a.uin, q, V5<T>{T(stagnant[0]), T(stagnant[1]), T(stagnant[2]), T(stagnant[3]), T(stagnant[4])});
}

// MarkerForce (cfd_ibm:442-480, 601-615): per marker, sample at offset `off` along the
// normal; pressure -p n A and viscous mu u_t / off A.  One thread per marker writes its six
// contributions (zero where the sample fails); one block then adds them in a fixed order
// (strided partial sums, then a fixed tree), so the FP64 total does not depend on scheduling.
template <class T>
__global__ __launch_bounds__(256) void marker3_eval_kernel(Args3 a, int n, const double* __restrict__ mk, double off,
                                                           double* __restrict__ contrib) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double* m = mk + 7 * q;
    const double sp[3] = {m[0] + off * m[3], m[1] + off * m[4], m[2] + off * m[5]};
    double c6[6] = {0, 0, 0, 0, 0, 0};
    P6<double> im;
    if (sample3<T>(a, (const T* const*)a.uin, sp, im) && isfinite(im.p) && isfinite(im.t) && isfinite(im.u) &&
        isfinite(im.v) && isfinite(im.w)) {
        const double us[3] = {im.u - 0.0, im.v - 0.0, im.w - 0.0};
        const double udn = us[0] * m[3] + us[1] * m[4] + us[2] * m[5];
        const double mu = a.visc_const ? a.mu : ref::sutherland(im.t);
        for (int d = 0; d < 3; d++) {
            const double ut = us[d] - udn * m[3 + d];
            c6[d] = -im.p * m[3 + d] * m[6];
            c6[3 + d] = mu * ut / off * m[6];
        }
    }
    for (int d = 0; d < 6; d++) contrib[6 * (long)q + d] = c6[d];
}
__global__ __launch_bounds__(256) void marker3_sum_kernel(const double* __restrict__ contrib, int n, double* out) {
    __shared__ double sh[256][6];
    const int t = threadIdx.x;
    double s6[6] = {0, 0, 0, 0, 0, 0};
    for (int q = t; q < n; q += 256)
        for (int d = 0; d < 6; d++) s6[d] += contrib[6 * (long)q + d];
    for (int d = 0; d < 6; d++) sh[t][d] = s6[d];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) for (int d = 0; d < 6; d++) sh[t][d] += sh[t + w][d];
        __syncthreads();
    }
    if (t == 0) for (int d = 0; d < 6; d++) out[d] = sh[0][d];
}

// Staircase wall force (cfd_ibm:309-362): pressure -p n A and wall shear mu u_t / (dn/2) A over
// every solid-fluid face of the raw state.  Its ghost cells hold the reference's values: each
// substage fills its input's ghosts and its update carries them into the output.  The ghost list
// is every solid cell with a fluid neighbour in raw order (buried cells have no fluid face), so
// lane l evaluates the six faces of ghost base + l and lane 0 adds the terms in cell order and
// face order x+, x-, y+, y-, z+, z-: the FP64 sum is the serial one.
template <class T>
__global__ void wall3_kernel(Args3 a, int n, const long* __restrict__ idx, double* out) {
    __shared__ double term[64][6][3];   // per face: pressure term, then the two tangential shear terms
    __shared__ unsigned char fluid[64][6];
    const int l = threadIdx.x;
    const T* const* U = (const T* const*)a.uin;
    const long st[3] = {1, a.NX, (long)a.NX * a.NY};
    double fp[3] = {0.0, 0.0, 0.0}, fv[3] = {0.0, 0.0, 0.0};
    for (int base = 0; base < n; base += 64) {
        const int q = base + l;
        if (q < n) {
            const long c = idx[q];
            const int jr = (int)((c / a.NX) % a.NY), kr = (int)(c / ((long)a.NX * a.NY)), ir = (int)(c % a.NX);
            const int rr[3] = {ir, jr, kr};
            double h[3];
            for (int d = 0; d < 3; d++) h[d] = a.spacing[d] ? a.spacing[d][rr[d] - NGH] : a.h[d];
            const double fa[3] = {h[1] * h[2], h[0] * h[2], h[0] * h[1]};
            for (int f = 0; f < 6; f++) {
                const int dim = f >> 1;
                const double sg = (f & 1) ? -1.0 : 1.0;
                const long nb = (f & 1) ? c - st[dim] : c + st[dim];
                const bool fl = !(a.solid && a.solid[nb]);
                fluid[l][f] = fl;
                if (!fl) continue;
                const V5<double> u = to_d(ld5(U, nb));
                term[l][f][0] = ref::pressure(u, a.gamma) * sg * fa[dim];
                const P6<double> pr = ref::cons_to_prim(u, a.gamma, a.r_gas);
                const double mu = a.visc_const ? a.mu : ref::sutherland(pr.t);
                const double vel[3] = {pr.u, pr.v, pr.w};
                for (int td = 0, t = 1; td < 3; td++)
                    if (td != dim) term[l][f][t++] = mu * vel[td] / (0.5 * h[dim]) * fa[dim];
            }
        }
        __syncthreads();
        if (l == 0) {
            for (int g = 0; g < 64 && base + g < n; g++)
                for (int f = 0; f < 6; f++) {
                    if (!fluid[g][f]) continue;
                    const int dim = f >> 1;
                    fp[dim] -= term[g][f][0];
                    for (int td = 0, t = 1; td < 3; td++)
                        if (td != dim) fv[td] += term[g][f][t++];
                }
        }
        __syncthreads();
    }
    if (l == 0)
        for (int d = 0; d < 3; d++) { out[d] = fp[d]; out[3 + d] = fv[d]; }
}

// ---------------------------------------------------------------------------
// Fused flux / viscous / RK kernel
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ V5<T> slope_at(const Args3& a, const T* const* U, long c, long s) {
    const V5<T> up = ld5(U, c + s), uc = ld5(U, c), un = ld5(U, c - s);
    const V5<T> f = ref::sub(up, uc), b = ref::sub(uc, un);
    const int lim = a.limiter;
    return {ref::limit(lim, f.r, b.r), ref::limit(lim, f.ru, b.ru), ref::limit(lim, f.rv, b.rv),
            ref::limit(lim, f.rw, b.rw), ref::limit(lim, f.re, b.re)};
}
// recon(): u + slope * frac along the face's axis.  The reference also adds the other
// axes' slopes times 0.0, which leaves every finite value unchanged.
template <class T>
__device__ __forceinline__ V5<T> recon(const V5<T>& u, const V5<T>& sl, T fr) {
    return ref::add(u, ref::mul(sl, fr));
}

// Flux through the + face of local cell `o` (index id_o along `ax`), flux_normal with sign +1.
template <class T, int FLUX, bool MUSCL>
__device__ __forceinline__ V5<T> plus_face(const Args3& a, const T* const* U, long o, int ax, int id_o, long s, int n) {
    const long oth = o + s;
    const bool other_local = id_o + 1 < n;
    const bool patch = (id_o + 1 == n) && !a.per[ax];
    V5<T> ul = ld5(U, o), ur = ld5(U, oth);
    if (MUSCL && !patch) {
        ul = recon(ul, slope_at(a, U, o, s), T(0.5));
        if (other_local) {
            ur = recon(ur, slope_at(a, U, oth, s), T(-0.5));
        } else if (a.seam_mode == 1) {
            const long w = oth - (long)n * s;
            ur = recon(ld5(U, w), slope_at(a, U, w, s), T(-0.5));
        }
    }
    return ref::num_flux(FLUX, ul, ur, T(a.gamma), ax);
}
// Flux through the - face of a cell on the low edge (id = 0), flux_normal with sign -1:
// -flux(other, owner).
template <class T, int FLUX, bool MUSCL>
__device__ __forceinline__ V5<T> minus_face(const Args3& a, const T* const* U, long o, int ax, long s, int n) {
    const long oth = o - s;
    const bool patch = !a.per[ax];
    V5<T> ul = ld5(U, o), ur = ld5(U, oth);
    if (MUSCL && !patch) {
        ul = recon(ul, slope_at(a, U, o, s), T(-0.5));
        if (a.seam_mode == 1) {
            const long w = oth + (long)n * s;
            ur = recon(ld5(U, w), slope_at(a, U, w, s), T(0.5));
        }
    }
    return ref::mul(ref::num_flux(FLUX, ur, ul, T(a.gamma), ax), T(-1));
}

template <class T>
__device__ __forceinline__ T d2_pre(T um, T u0, T up, T ihm, T ihp, T ihs) {
    return ((up - u0) * ihp - (u0 - um) * ihm) * ihs;   // cfd_solver:449-453
}

// Viscous terms (cfd_solver:463-560), body force, solid-rhs zero (cfd_ibm:293-301), then
// the RK substep (integrate.rs:75-101) of raw cell q; rhs holds the inviscid divergence.
template <class T>
__device__ __forceinline__ void finish_cell(const Args3& a, const T* const* U, const T* const* U0, T* const* O, long q,
                                            const int id[3], bool interior, V5<T> rhs, T dt) {
    const long st[3] = {1, a.NX, (long)a.NX * a.NY};
    const T g = T(a.gamma), R = T(a.r_gas);
    const V5<T> uc = ld5(U, q);
    if (interior) {
        const bool solid_c = a.solid && a.solid[q];
        if (a.viscous && !solid_c) {
            const P6<T> ps = ref::cons_to_prim(uc, g, R);
            const T mu = a.visc_const ? T(a.mu) : ref::sutherland(ps.t);
            const T kap = ref::conductivity(mu, g, R);
            const long nb[6] = {q + 1, q - 1, q + st[1], q - st[1], q + st[2], q - st[2]};
            T qv[6][4];
            for (int t = 0; t < 6; t++) {
                if (a.solid && a.solid[nb[t]]) {
                    qv[t][0] = T(0); qv[t][1] = T(0); qv[t][2] = T(0); qv[t][3] = ps.t;
                } else {
                    const P6<T> p = ref::cons_to_prim(ld5(U, nb[t]), g, R);
                    qv[t][0] = p.u; qv[t][1] = p.v; qv[t][2] = p.w; qv[t][3] = p.t;
                }
            }
            const T vel[4] = {ps.u, ps.v, ps.w, ps.t};
            const T dxi2 = T(a.dxi2);
            const T yim = T(a.yw[id[1]]), yip = T(a.yw[a.ny + id[1]]), yis = T(a.yw[2 * a.ny + id[1]]);
            const T zim = T(a.zw[id[2]]), zip = T(a.zw[a.nz + id[2]]), zis = T(a.zw[2 * a.nz + id[2]]);
            T lap[4];
            for (int m = 0; m < 4; m++) {
                const T d2x = (qv[0][m] - T(2) * vel[m] + qv[1][m]) * dxi2;
                const T d2y = d2_pre(qv[3][m], vel[m], qv[2][m], yim, yip, yis);
                const T d2z = d2_pre(qv[5][m], vel[m], qv[4][m], zim, zip, zis);
                lap[m] = d2x + d2y + d2z;
            }
            const T fx = mu * lap[0], fy = mu * lap[1], fz = mu * lap[2];
            const T en = vel[0] * fx + vel[1] * fy + vel[2] * fz + kap * lap[3];
            rhs = ref::add(rhs, V5<T>{T(0), fx, fy, fz, en});
        }
        if (a.body_force) {   // poiseuille_channel_validation.rs:105-120
            const T gx = T(a.gx);
            rhs.ru += uc.r * gx;
            rhs.re += gx * uc.ru;
        }
        if (solid_c) rhs = V5<T>{T(0), T(0), T(0), T(0), T(0)};
    }
    if (a.rhs_only) {   // the substage's rhs column (zero on ghost and solid cells), for checkpoints
        st5(O, q, rhs);
        return;
    }
    const V5<T> inc = ref::add(uc, ref::mul(rhs, dt));
    V5<T> o;
    if (a.euler || a.stage == 0) o = inc;
    else if (a.stage == 1) o = ref::add(ref::mul(ld5(U0, q), T(0.75)), ref::mul(inc, T(0.25)));
    else o = ref::add(ref::mul(ld5(U0, q), T(1.0 / 3.0)), ref::mul(inc, T(2.0 / 3.0)));
    st5(O, q, o);
}

// Face area over cell volume of interior cell (., j, k) per axis (crates/cfd_solver/src/lib.rs:
// 418-424: area dy dz, dx dz, dx dy over the volume dx dy dz of the cell's own spacings; a face
// shared along an axis has the same area on both sides, so each cell scales what it receives by
// its own ratios).  Uniform spacing: the precomputed ratios.
__device__ __forceinline__ void cell_aov(const Args3& a, int j, int k, double o[3]) {
    if (!a.spacing[1] && !a.spacing[2]) {
        o[0] = a.aov[0]; o[1] = a.aov[1]; o[2] = a.aov[2];
        return;
    }
    const double dx = a.h[0], dy = a.spacing[1] ? a.spacing[1][j] : a.h[1], dz = a.spacing[2] ? a.spacing[2][k] : a.h[2];
    const double vol = dx * dy * dz;
    o[0] = (dy * dz) / vol; o[1] = (dx * dz) / vol; o[2] = (dx * dy) / vol;
}

// One thread per raw cell: each interior cell evaluates its own faces and those of its
// lower neighbours, so every interior face is evaluated twice, by the two cells it
// separates, with the same operands.  (An x-y tiled variant that evaluates each face once
// into LDS, marching 1 to 4 z levels per block, measured 1.35-1.4x slower: the FP64 face
// work is latency-bound and the barrier-separated phases leave too few waves in flight.)
template <class T, int FLUX, bool MUSCL>
__global__ __launch_bounds__(256) void stage3_kernel(Args3 a) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const double dtd = dt_of(a);
    if (q == 0 && a.stage == 0 && !a.rhs_only) a.clock->dt = dtd;
    if (q >= a.nraw) return;
    const T* const* U = (const T* const*)a.uin;
    const int i = (int)(q % a.NX);
    const long jk = q / a.NX;
    const int j = (int)(jk % a.NY), k = (int)(jk / a.NY);
    const int id[3] = {i - NGH, j - NGH, k - NGH};
    const int n[3] = {a.nx, a.ny, a.nz};
    const long st[3] = {1, a.NX, (long)a.NX * a.NY};
    V5<T> rhs = {T(0), T(0), T(0), T(0), T(0)};
    const bool interior = id[0] >= 0 && id[0] < n[0] && id[1] >= 0 && id[1] < n[1] && id[2] >= 0 && id[2] < n[2];
    if (interior) {
        double aov[3];
        cell_aov(a, id[1], id[2], aov);
        // + faces of the lower neighbours, in the sweep's visiting order (z, y, x)
        for (int ax = 2; ax >= 0; --ax) {
            if (id[ax] > 0) {
                const V5<T> f = plus_face<T, FLUX, MUSCL>(a, U, q - st[ax], ax, id[ax] - 1, st[ax], n[ax]);
                rhs = ref::add(rhs, ref::mul(f, T(aov[ax])));
            }
        }
        // own faces: + face, then the - face on a low edge, per axis
        for (int ax = 0; ax < 3; ++ax) {
            const T inv = T(aov[ax]);
            rhs = ref::sub(rhs, ref::mul(plus_face<T, FLUX, MUSCL>(a, U, q, ax, id[ax], st[ax], n[ax]), inv));
            if (id[ax] == 0) rhs = ref::sub(rhs, ref::mul(minus_face<T, FLUX, MUSCL>(a, U, q, ax, st[ax], n[ax]), inv));
        }
    }
    finish_cell<T>(a, U, (const T* const*)a.u0, (T* const*)a.uout, q, id, interior, rhs, T(dtd));
    if (q == 0 && (a.euler || a.stage == 2) && !a.rhs_only) {
        a.clock->time += dtd;
        a.clock->step += 1;
    }
}

// ---------------------------------------------------------------------------
// CFL scan (crates/cfd_solver/src/lib.rs:626-677): per-block maxima, then one block.
// Cells with rho <= 0 are skipped, as in the reference; they and non-finite lambdas
// are counted for the non-physical flag.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void cfl3_kernel(Args3 a, double* part) {
    __shared__ double red[4];
    double m = 0.0;
    long long bad = 0;
    const T* const* U = (const T* const*)a.uin;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < a.ncell; c += (long)gridDim.x * blockDim.x) {
        const int i = (int)(c % a.nx);
        const long jk = c / a.nx;
        const int j = (int)(jk % a.ny), k = (int)(jk / a.ny);
        const long q = (i + NGH) + (long)a.NX * ((j + NGH) + (long)a.NY * (k + NGH));
        const V5<double> u = to_d(ld5(U, q));
        if (!(u.r > 0.0)) { bad++; continue; }
        double v[3];
        ref::velocity(u, v);
        const double s = ref::sound(u, a.gamma);
        const double dx = a.h[0], dy = a.spacing[1] ? a.spacing[1][j] : a.h[1], dz = a.spacing[2] ? a.spacing[2][k] : a.h[2];
        double lam = (fabs(v[0]) + s) / dx + (fabs(v[1]) + s) / dy + (fabs(v[2]) + s) / dz;
        if (a.viscous) {
            const P6<double> p = ref::cons_to_prim(u, a.gamma, a.r_gas);
            const double nu = (a.visc_const ? a.mu : ref::sutherland(p.t)) / u.r;
            lam += 2.0 * nu * (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
        }
        if (lam == lam) m = fmax(m, lam);
        else bad++;
    }
    m = wave_max(m);
    bad = wave_sum_ll(bad);
    if (lane_id() == 0) {
        red[threadIdx.x >> 6] = m;
        if (bad) atomicAdd((unsigned long long*)&a.clock->bad, (unsigned long long)bad);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
        part[blockIdx.x] = fmax(m, red[0]);
    }
}
__global__ __launch_bounds__(1024) void cfl3_final_kernel(const double* __restrict__ part, int n, double* target) {
    __shared__ double red[16];
    double m = 0.0;
    for (int b = threadIdx.x; b < n; b += blockDim.x) m = fmax(m, part[b]);
    m = wave_max(m);
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
        *target = m;
    }
}

template <class T>
__global__ __launch_bounds__(256) void sponge3_kernel(Args3 a, Sponge3 s) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncell) return;
    const int i = (int)(c % a.nx);
    const long jk = c / a.nx;
    const int j = (int)(jk % a.ny), k = (int)(jk / a.ny);
    const double cx = a.lo[0] + ((double)i + 0.5) * a.h[0];
    if (cx < s.x0) return;
    const double dt = a.clock->dt;
    const double inv_w = 1.0 / (s.x1 - s.x0);
    double sc = (cx - s.x0) * inv_w;
    sc = sc < 0.0 ? 0.0 : (sc > 1.0 ? 1.0 : sc);
    const double f = fmin(s.sigma * sc * sc * dt, 1.0);
    V5<double> tg;
    if (s.uniform) {
        tg = {s.target[0], s.target[1], s.target[2], s.target[3], s.target[4]};
    } else {
        const double cy = a.cen[1] ? a.cen[1][j] : a.lo[1] + ((double)j + 0.5) * a.h[1];
        const double uy = 4.0 * s.umax * cy * (s.height - cy) / (s.height * s.height);
        P6<double> pv = {s.rho, uy, 0.0, 0.0, s.p, s.p / (s.rho * 287.058)};
        tg = ref::prim_to_cons(pv, a.gamma);
    }
    const long q = (i + NGH) + (long)a.NX * ((j + NGH) + (long)a.NY * (k + NGH));
    T* const* W = (T* const*)a.uout;
    V5<double> u = to_d(ld5((const T* const*)a.uout, q));
    u.r += f * (tg.r - u.r);
    u.ru += f * (tg.ru - u.ru);
    u.rv += f * (tg.rv - u.rv);
    u.rw += f * (tg.rw - u.rw);
    u.re += f * (tg.re - u.re);
    st5(W, q, from_d<T>(u));
}

// FP64 interior sums of the five components: per-block partials in a fixed order,
// then one block combines them in block order, so the totals are deterministic.
template <class T>
__global__ __launch_bounds__(256) void sums3_kernel(Args3 a, double* part) {
    __shared__ double sh[256][5];
    const T* const* U = (const T* const*)a.uin;
    double s5[5] = {0, 0, 0, 0, 0};
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < a.ncell; c += (long)gridDim.x * blockDim.x) {
        const int i = (int)(c % a.nx);
        const long jk = c / a.nx;
        const int j = (int)(jk % a.ny), k = (int)(jk / a.ny);
        const long q = (i + NGH) + (long)a.NX * ((j + NGH) + (long)a.NY * (k + NGH));
        if (a.spacing[1] || a.spacing[2]) {   // stretched: each cell weighed by its volume
            const double vol = a.h[0] * (a.spacing[1] ? a.spacing[1][j] : a.h[1]) * (a.spacing[2] ? a.spacing[2][k] : a.h[2]);
            for (int m = 0; m < 5; m++) s5[m] += (double)U[m][q] * vol;
        } else {
            for (int m = 0; m < 5; m++) s5[m] += (double)U[m][q];
        }
    }
    for (int m = 0; m < 5; m++) sh[threadIdx.x][m] = s5[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) for (int m = 0; m < 5; m++) sh[threadIdx.x][m] += sh[threadIdx.x + w][m];
        __syncthreads();
    }
    if (threadIdx.x == 0) for (int m = 0; m < 5; m++) part[5 * blockIdx.x + m] = sh[0][m];
}
__global__ void sums3_final_kernel(const double* part, int nb, double vol, double* out) {
    if (threadIdx.x != 0) return;
    double s[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < nb; b++) for (int m = 0; m < 5; m++) s[m] += part[5 * b + m];
    for (int m = 0; m < 5; m++) out[m] = s[m] * vol;
}

__global__ void clock3_set_kernel(DevClock* clk, double dt, double time, long long step) {
    clk->dt = dt; clk->time = time; clk->step = step; clk->bad = 0; clk->lam[0] = 0.0; clk->lam[1] = 0.0;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {
inline unsigned blocks_for(long n, int b) { return (unsigned)((n + b - 1) / b); }

template <class T, int FLUX>
void stage3_dispatch(const Args3& a, hipStream_t st) {
    if (a.muscl) stage3_kernel<T, FLUX, true><<<blocks_for(a.nraw, 256), 256, 0, st>>>(a);
    else stage3_kernel<T, FLUX, false><<<blocks_for(a.nraw, 256), 256, 0, st>>>(a);
}
template <class T>
void stage3_launch(const Args3& a, hipStream_t st) {
    switch (a.flux) {
    case FLUX_RUSANOV: stage3_dispatch<T, FLUX_RUSANOV>(a, st); break;
    case FLUX_HLLC: stage3_dispatch<T, FLUX_HLLC>(a, st); break;
    default: stage3_dispatch<T, FLUX_HLLC_TEXTBOOK>(a, st); break;
    }
}
}  // namespace

hipError_t launch3_fill(const Args3& a, int esz, hipStream_t st) {
    for (int d = 0; d < 3; d++) {
        if (!a.per[d]) continue;
        if (esz == 8) halo3_kernel<double><<<blocks_for(a.nraw, 256), 256, 0, st>>>(a, d);
        else halo3_kernel<float><<<blocks_for(a.nraw, 256), 256, 0, st>>>(a, d);
    }
    long mf = 0;
    const long f[3] = {(long)a.ny * a.nz, (long)a.nx * a.nz, (long)a.nx * a.ny};
    for (int s = 0; s < 6; s++)
        if (a.bc_kind[s] != BC_PERIODIC && f[s / 2] > mf) mf = f[s / 2];
    if (mf > 0) {
        dim3 grid(blocks_for(mf, 256), 6);
        if (esz == 8) bc3_kernel<double><<<grid, 256, 0, st>>>(a);
        else bc3_kernel<float><<<grid, 256, 0, st>>>(a);
    }
    return hipGetLastError();
}

hipError_t launch3_mirror(const Args3& a, int esz, int n, const long* idx, const double* img, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (esz == 8) mirror3_kernel<double><<<(n + 127) / 128, 128, 0, st>>>(a, n, idx, img);
    else mirror3_kernel<float><<<(n + 127) / 128, 128, 0, st>>>(a, n, idx, img);
    return hipGetLastError();
}

hipError_t launch3_flip(const Args3& a, int esz, const double* stagnant, hipStream_t st) {
    if (esz == 8) flip3_kernel<double><<<blocks_for(a.nraw, 256), 256, 0, st>>>(a, stagnant);
    else flip3_kernel<float><<<blocks_for(a.nraw, 256), 256, 0, st>>>(a, stagnant);
    return hipGetLastError();
}

hipError_t launch3_stage(const Args3& a, int esz, hipStream_t st) {
    if (esz == 8) stage3_launch<double>(a, st);
    else stage3_launch<float>(a, st);
    return hipGetLastError();
}

int cfl3_blocks(long ncell) {
    const long nb = (ncell + 255) / 256;
    return (int)(nb < 1024 ? nb : 1024);
}

hipError_t launch3_cfl(const Args3& a, int esz, double* part, double* target, hipStream_t st) {
    const int nb = cfl3_blocks(a.ncell);
    if (esz == 8) cfl3_kernel<double><<<nb, 256, 0, st>>>(a, part);
    else cfl3_kernel<float><<<nb, 256, 0, st>>>(a, part);
    cfl3_final_kernel<<<1, 1024, 0, st>>>(part, nb, target);
    return hipGetLastError();
}

hipError_t launch3_sponge(const Args3& a, int esz, const Sponge3& s, hipStream_t st) {
    if (esz == 8) sponge3_kernel<double><<<blocks_for(a.ncell, 256), 256, 0, st>>>(a, s);
    else sponge3_kernel<float><<<blocks_for(a.ncell, 256), 256, 0, st>>>(a, s);
    return hipGetLastError();
}

hipError_t launch3_markers(const Args3& a, int esz, int n, const double* mk, double off, double* contrib, double* out,
                           hipStream_t st) {
    if (n > 0) {
        if (esz == 8) marker3_eval_kernel<double><<<(n + 255) / 256, 256, 0, st>>>(a, n, mk, off, contrib);
        else marker3_eval_kernel<float><<<(n + 255) / 256, 256, 0, st>>>(a, n, mk, off, contrib);
    }
    marker3_sum_kernel<<<1, 256, 0, st>>>(contrib, n, out);
    return hipGetLastError();
}

hipError_t launch3_wall(const Args3& a, int esz, int n, const long* idx, double* out, hipStream_t st) {
    if (esz == 8) wall3_kernel<double><<<1, 64, 0, st>>>(a, n, idx, out);
    else wall3_kernel<float><<<1, 64, 0, st>>>(a, n, idx, out);
    return hipGetLastError();
}

hipError_t launch3_sums(const Args3& a, int esz, double* part, int nb, double vol, double* out, hipStream_t st) {
    if (esz == 8) sums3_kernel<double><<<nb, 256, 0, st>>>(a, part);
    else sums3_kernel<float><<<nb, 256, 0, st>>>(a, part);
    sums3_final_kernel<<<1, 64, 0, st>>>(part, nb, vol, out);
    return hipGetLastError();
}

hipError_t launch3_clock_set(DevClock* clk, double dt, double time, long long step, hipStream_t st) {
    clock3_set_kernel<<<1, 1, 0, st>>>(clk, dt, time, step);
    return hipGetLastError();
}

}  // namespace efvm
