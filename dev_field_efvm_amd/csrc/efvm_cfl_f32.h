// efvm_cfl_f32.h -- the per-cell CFL spectral radius of the FP32 stage kernels, shared by
// every translation unit that fuses the CFL scan into a stage (the MUSCL stage kernels and
// the learned-reconstruction stage kernels), so that all of them produce the same value.
#pragma once
#include "efvm_stage.h"
#include "efvm_fast.h"

namespace efvm {

// CFL spectral radius of one FP32 cell (crates/cfd_solver/src/lib.rs:653-666, nz = 1), with
// the operation order fixed by explicit fmaf so that the fused epilogues and the standalone
// FP32 scan produce the same value: a run restarted from a checkpoint takes the same dt.
// Returns false for a non-physical cell (rho <= 0, p <= 0 or a non-finite lambda).
template <bool VISC>
__device__ __forceinline__ bool lam_cell_f32(const StageArgs& a, float o0, float o1, float o2, float o3, float& l) {
    const float g = (float)a.gamma, gm1 = (float)(a.gamma - 1.0);
    const float idx = (float)(1.0 / a.dx), idy = (float)(1.0 / a.dy), idz = (float)(1.0 / a.dz);
    const float ir = fast::rcp(o0);
    const float vu = fabsf(o1 * ir), vv = fabsf(o2 * ir);
    const float q2 = fmaf(o1, o1, o2 * o2);
    const float p = gm1 * fmaf(-0.5f * q2, ir, o3);
    const float snd = fast::fsqrt(g * p * ir);
    l = fmaf(vu + snd, idx, fmaf(vv + snd, idy, snd * idz));
    if (VISC) {
        const float vsum = (float)(1.0 / (a.dx * a.dx) + 1.0 / (a.dy * a.dy) + 1.0 / (a.dz * a.dz));
        const float t = p * ir * (float)(1.0 / a.r_gas);
        const float mu = a.visc_const ? (float)a.mu : (float)ref::sutherland((double)t);
        l = fmaf(2.f * mu * ir, vsum, l);
    }
    // positive density and pressure and a finite lambda, tested on the bit patterns (a unit
    // compiled assuming no NaNs could fold a float compare): a positive float is a
    // positive integer, and a finite non-negative lambda lies below the pattern of +inf (NaNs
    // and negative values compare above it as unsigned integers)
    return (int)__builtin_bit_cast(unsigned, o0) > 0 && (int)__builtin_bit_cast(unsigned, p) > 0 &&
           __builtin_bit_cast(unsigned, l) < 0x7f800000u;
}

}  // namespace efvm
