// efvm_stage3d.h -- argument blocks and host launchers of the 3-D structured-mesh
// kernels (efvm_stage3d.hip), shared with the host API (efvm_api3d.cpp).
#pragma once
#include "efvm_stage.h"

namespace efvm {

constexpr int NGH = 2;   // ghost layers per side

struct Args3 {
    int nx, ny, nz;        // interior cells
    int NX, NY, NZ;        // raw (padded) extents, n + 2 NGH
    long nraw, ncell;
    double lo[3], h[3];
    const double* spacing[3];   // per-axis cell widths (interior cells) of a stretched axis, or null: h
    const double* cen[3];       // per-axis cell centroids (interior cells) of a stretched axis, or null
    double aov[3];         // face area / cell volume per axis
    int per[3];
    int seam_mode;
    int bc_kind[6];        // XLo XHi YLo YHi ZLo ZHi
    double bc_par[6][5];
    double bc_cons[6][5];  // SupersonicInflow ghost state (prim_to_cons of its parameters)
    const double* tab[6];  // FromOutside ghost tables: per face (rho, rho_u, rho_v, rho_w, rho_E, given), or null
    double gamma, r_gas;
    int flux, muscl, limiter;
    int viscous, visc_const;
    double mu;
    double dxi2;
    const double* yw;      // [3 ny]: 1/hm, 1/hp, 2/(hm+hp) per interior row (cfd_solver:483-501)
    const double* zw;      // [3 nz]: the same per layer
    int body_force;
    double gx;
    const uint8_t* solid;  // raw-indexed mask, or null
    int euler, stage, parity;
    int rhs_only;          // 1: write the right-hand side to uout instead of the RK update (no clock change)
    double cfl, fixed_dt;
    DevClock* clock;
    void* uin[5];
    void* u0[5];
    void* uout[5];
};

// Sponge after a completed step (examples/sphere_drag_validation/main.rs:119-160 with a
// uniform target; examples/cylinder_drag_validation/main.rs:186-229 with the parabolic one).
struct Sponge3 {
    double x0, x1, sigma;
    int uniform;
    double target[5];                 // uniform: conserved target state
    double umax, height, rho, p;      // parabolic: u(y) = 4 umax y (h - y) / h^2
};
hipError_t launch3_fill(const Args3& a, int esz, hipStream_t st);
hipError_t launch3_mirror(const Args3& a, int esz, int n, const long* idx, const double* img, hipStream_t st);
hipError_t launch3_flip(const Args3& a, int esz, const double* stagnant, hipStream_t st);
hipError_t launch3_stage(const Args3& a, int esz, hipStream_t st);
int cfl3_blocks(long ncell);
hipError_t launch3_cfl(const Args3& a, int esz, double* part, double* target, hipStream_t st);
hipError_t launch3_sponge(const Args3& a, int esz, const Sponge3& s, hipStream_t st);
hipError_t launch3_wall(const Args3& a, int esz, int n, const long* idx, double* out, hipStream_t st);
hipError_t launch3_markers(const Args3& a, int esz, int n, const double* mk, double off, double* contrib, double* out,
                           hipStream_t st);
hipError_t launch3_sums(const Args3& a, int esz, double* part, int nb, double vol, double* out, hipStream_t st);
hipError_t launch3_clock_set(DevClock* clk, double dt, double time, long long step, hipStream_t st);

}  // namespace efvm
