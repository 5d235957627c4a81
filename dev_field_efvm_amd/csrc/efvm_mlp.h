// efvm_mlp.h -- the learned face reconstruction: network layout, feature
// definition and the device argument blocks shared by the inference, training
// and sample-generation kernels (efvm_mlp.hip), the learned stage kernel
// (efvm_stage_f32.hip) and the host (efvm_api.cpp).
//
// The network maps the 20 features of one face side (a 5-cell stencil of the
// primitives rho, u_n, u_t, p, oriented toward the face) to the normalised face
// offsets y (4 values):
//     y = (x[+1] - x[-1]) / 4 + phi * (L x + bL),
//     phi = sigmoid(W3 gelu(W2 gelu(W1 x + b1) + b2) + b3)      (per variable),
// face value (oriented) = w_centre + sigma * y.  The first term is the unlimited
// central slope, L x + bL a learned high-order correction, and the MLP (20-64-64-4,
// tanh-form GELU) a learned limiter that gates the correction between fully applied
// (phi = 1) and off (phi = 0): it can limit, never amplify.  Inference runs the MLP
// on bf16 matrix cores with FP32 accumulation; the correction, the gate and the
// offsets stay FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efvm {
namespace mlp {

constexpr int NIN = 20, NH = 64, NOUT = 4;
// flat FP32 parameter vector (row-major matrices, [out][in])
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NOUT * NH;
constexpr int OFF_L = OFF_B3 + NOUT;
constexpr int OFF_BL = OFF_L + NOUT * NIN;
constexpr int NPARAM = OFF_BL + NOUT;   // 5848
constexpr int TB = 64;                  // samples per training block (one wave)
constexpr int STAT_SLOTS = 8192;        // waves of the largest face-kernel launch
constexpr int LIGHT_OUT = 58;            // output columns per tile of the fused bypass stage (a
                                        // wave of 64 columns less the +-3 of the x-face stencil)
constexpr int LIGHT_ROWS = 3;            // rows per tile (a workgroup of LIGHT_ROWS + 1 waves; 3 keeps a
                                        // 512^2 stage's workgroups resident in one round)

// features: x[4k+v], k = 0..4 <-> stencil cells -2..2 (k = 3 is the neighbour across
// the face), v = rho, u_n, u_t, p.  x = (w_k - w_0) / sigma_v for k != 2 and
// x[8+v] = sigma_v / (sigma_v + X0C scale_v); sigma_v = max(max_k |w_k - w_0|, FLOOR scale_v),
// scale = (rho_0, a_0, a_0, p_0).
constexpr float FLOOR = 1.0e-4f;
constexpr float X0C = 0.1f;

// Training loss: mean over samples and variables of w (y - y*)^2 with the weight
//     w_v = min((r_v / WREF)^2, WMAX),   r_v = sigma_v / scale_v = X0C x0 / (1 - x0)
// recovered from the centre feature x0 = x[8+v].  Normalised offsets make every face
// side count alike; weighting by sigma^2 turns the loss into the squared error of the
// face values themselves (what the solution sees), capped so shocks do not dominate.
constexpr float WREF = 0.01f;
constexpr float WMAX = 100.f;
__host__ __device__ inline float sample_weight(float x0) {
    const float r = X0C * x0 / (1.f - x0);
    const float w = (r / WREF) * (r / WREF);
    return w < WMAX ? w : WMAX;
}
constexpr float GELU_C = 0.7978845608028654f;   // sqrt(2/pi)
constexpr float GELU_A = 0.044715f;

// Inference weights in the lane order of the bf16 16x16x32 matrix-core operands.
// Layer 1 takes K = 32 inputs: the 20 features, a constant 1 (feature 20, whose
// weight column is b1) and zeros; b3 is added with the FP32 offsets.
// Layer l computes Z^T (hidden x 16 face sides) = W (hidden x K) . X^T (K x 16):
// the A operand of lane `ln`, element e is W[row 16t + ln%16][k = 8(ln/16) + e].
// Between layers the K index is permuted so that the accumulator a lane holds is
// exactly the B operand element it needs: hidden unit
//     h(s, g, e) = 16 (2 s + e / 4) + 4 g + e % 4      (K step s, lane group g)
// Layer 3 replicates the 4 output rows over the 4 row groups, so every lane
// group receives the outputs of its own 16 face sides.
struct Packed {
    uint32_t w1[4][64][4];       // [M tile][lane][8 bf16]
    uint32_t w2[4][2][64][4];    // [M tile][K step][lane][8 bf16]
    uint32_t w3[2][64][4];       // [K step][lane][8 bf16]
    float b1[NH], b2[NH], b3[NOUT];
    float L[NOUT][NIN], bL[NOUT];
    int own4;                    // L only couples each variable to its own 4 neighbour cells and
                                 // bL = 0 (the cubic-exact fit): the correction is 4 taps, not 20,
                                 // and with the gate open the face value is a fixed linear stencil
};

// one launch of the face-state kernel
struct FaceArgs {
    int nx, ny;
    int n_chunks;                // ceil(nx / 64)
    int n_tasks;                 // 2 * ny * n_chunks (x faces, then y faces)
    float gamma;
    const float* U[4];           // conserved planes of the stage input
    float* F;                    // out_mode 0: 16 planes ((axis * 2 + side) * 4 + var) * nx*ny of face
                                 // primitives; out_mode 1: 8 planes (axis * 4 + var) * nx*ny of the
                                 // Riemann flux through the east / north face of each cell
    int out_mode;
    int flux;                    // out_mode 1: 0 Rusanov, 1 HLLC (Thornber), 2 HLLC
    const Packed* w;
    unsigned long long* stats;   // per-wave counters, 3 per wave slot (STAT_SLOTS slots), accumulated:
                                 // [faces that fell back to MUSCL, faces evaluated, face sides
                                 // whose gate the network evaluated]; each wave of a launch owns
                                 // its slot, so no two waves touch one counter
    int bypass;                  // 1: the network's gate is evaluated only for face sides whose
                                 // ungated value fails the TVD bound with half the relaxation;
                                 // elsewhere the gate is fully open (phi = 1)
    unsigned* list;              // deferred faces of the bypass passes: axis * nx*ny + row * nx + column
    unsigned* n_list;            // their count (device)
    int own4;                    // host copy of w->own4 (selects the light pass's specialisation)
    int n_lam_slots;             // fused bypass stage: the per-wave CFL maxima pending_kernel folds
                                 // into the clock on the last stage (0: none)
    float taps[4][4];            // own4: the open-gate value's taps on the stencil cells -2, -1, +1, +2
                                 // per variable, L's own-variable weights with the central slope's
                                 // -1/4 and +1/4 folded into the inner two (uniform, so they sit in
                                 // scalar registers)
};

// the taps of FaceArgs::taps from the correction's weights L[v][f] (feature f = 4 k + v)
inline void light_taps(const float L[NOUT][NIN], float taps[4][4]) {
    for (int v = 0; v < 4; v++) {
        taps[0][v] = L[v][v];
        taps[1][v] = L[v][4 + v] - 0.25f;
        taps[2][v] = L[v][12 + v] + 0.25f;
        taps[3][v] = L[v][16 + v];
    }
}

// one training minibatch (loss and gradient)
struct TrainArgs {
    const float* params;          // NPARAM
    const float* X;               // samples x 20
    const float* Y;               // samples x 4
    long n_data;                  // samples available
    long n;                       // samples in this batch
    int random;                   // 1: batch drawn uniformly from the data (counter hash of seed, step, slot)
    unsigned long long seed, step;
    float scale;                  // 2 / (n * NOUT): d loss / d y
    float* partial;               // blocks x NPARAM
    double* loss_partial;         // blocks
};

// training samples from a fine periodic state box-filtered by f
struct SampleArgs {
    int nxf, nyf, f, nc_x, nc_y;  // fine size, factor, coarse size
    double gamma;
    const void* U[4];             // fine conserved planes
    int esz;                      // 4 or 8
    const void* Uc[4];            // coarse solver's planes (the features' cells), or null: box filter
    int esz_c;
    double* cells;                // 4 planes of coarse cells
    double* faces;                // 2 axes x 4 planes: east / north face line averages
    float* X;                     // output rows (samples x 20)
    float* Y;                     // output rows (samples x 4)
};

}  // namespace mlp
}  // namespace efvm
