// Learned warm start: a message-passing network over the cloth edge graph that
// predicts each vertex's XPBD position correction for the coming substep.
// No reference counterpart; it sits between predict and the projection loop of
// the reference substep (crates/cloth_core/src/simulation/dispatch.rs:137 -> 169).
//
// Network (hidden width H = 128, input width 32, R message-passing rounds):
//   encoder   h0      = W_e2 relu(W_e1 f + b_e1) + b_e2
//   round r   h_{r+1} = h_r + W_r2 relu(W_r1 [h_r, mean_{j in N(i)} h_r[j]] + b_r1) + b_r2
//   decoder   y       = W_d2 relu(W_d1 h_R + b_d1) + b_d2          (16 outputs, the first 3 used)
// Dense layers run on MFMA (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulation); activations are stored in BF16, gradients accumulate in FP32.
// The correction applied to a free vertex is  beta * D_prev + s * y[0:3],  where
// D_prev is the correction of the previous substep, beta its weight (the history
// weight) and s = max(|g|, 1 m/s^2) dt^2.  The previous correction also reaches the
// next substep through the velocity, so beta = 1 feeds slowly corrected error modes
// back with a gain above one; beta <= 0.5 keeps that loop contracting.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tc {
namespace ws {

constexpr int H = 128;        // hidden width
constexpr int F = 32;         // input features per vertex
constexpr int O = 16;         // decoder output width
constexpr int TR = 64;        // rows per tile
constexpr int MAX_LAYERS = 16;
// Rows of the buffers a message-passing round writes (h, and the training saves x_r, a_r):
// whole tiles, so that the round kernel stores without bounds checks.
constexpr size_t padded_rows(size_t rows) { return (rows + TR - 1) / TR * TR; }

typedef __bf16 bf16;

// One dense layer inside the flat parameter vector: W (K x N, row-major) then b (N).
struct Layer { uint32_t off, K, N; };

struct NetLayout {
    uint32_t rounds = 3;
    uint32_t nlayers = 0;
    Layer L[MAX_LAYERS];
    uint32_t nparams = 0;
    // layer indices
    uint32_t enc1() const { return 0; }
    uint32_t enc2() const { return 1; }
    uint32_t r1(uint32_t r) const { return 2 + 2 * r; }
    uint32_t r2(uint32_t r) const { return 3 + 2 * r; }
    uint32_t dec1() const { return 2 + 2 * rounds; }
    uint32_t dec2() const { return 3 + 2 * rounds; }
};
NetLayout make_layout(uint32_t rounds);

// Device copies of the weights: fp32 master (theta), bf16 row-major (w) and bf16
// transposed (wt) at the same offsets as theta.
struct NetDev {
    const float *theta;
    const bf16 *w;
    const bf16 *wt;
};

// Feature inputs of one cloth.
struct FeatArgs {
    const float4 *P;        // predicted positions + inverse mass
    const float4 *D;        // previous correction (xyz)
    const uint32_t *d_off, *d_nbr; const float *d_rest;  // distance-constraint incidence (CSR)
    const uint32_t *b_off, *b_nbr; const float *b_rest;  // bending-pair incidence (CSR)
    const void *cols; uint32_t ncols;                    // tc::Collider primitives
    const float4 *X;        // committed positions (w = thickness)
    uint32_t nv;
    float s;                // max(|g|, 1) dt^2
    float lscale;           // collider distance scale (mean edge length)
    float beta;             // history weight: strains are taken at prediction + beta * D
};

// Graph of the message-passing rounds (stretch edges, both directions); max_degree is the
// largest row length (it picks how many neighbours a round prefetches per row).
struct Graph { const uint32_t *off, *nbr; uint32_t max_degree; };
inline uint32_t max_degree(const uint32_t *off, uint32_t rows) {
    uint32_t m = 0;
    for (uint32_t i = 0; i < rows; ++i) m = off[i + 1] - off[i] > m ? off[i + 1] - off[i] : m;
    return m;
}

hipError_t launch_features(hipStream_t st, const FeatArgs &a, bf16 *feat);
hipError_t launch_copy_pred(hipStream_t st, const float4 *P, float4 *Ppred, uint32_t nv);
// Applies beta * D_prev + s*y (y == nullptr: history only) to the free vertices of P; keeps P in Ppred.
hipError_t launch_apply(hipStream_t st, float4 *P, float4 *Ppred, const float4 *D, const float *y, float s,
                        float beta, uint32_t nv);
// D = P - Ppred on free vertices (0 on pinned ones); with T != nullptr also writes the
// training target T = ((P - Ppred - beta * D_old) / s, free ? 1 : 0) before D is replaced.
hipError_t launch_update_prev(hipStream_t st, const float4 *P, const float4 *Ppred, float4 *D, float4 *T, float s,
                              float beta, uint32_t nv);

// Forward of the whole network (inference or training).  Buffers (rows = nv):
//   feat rows x F;  h[0..R] padded_rows(rows) x H (h[r] for training, two ping-pong buffers
//   otherwise);
//   y rows x O (fp32).  Training saves (may be null for inference):
//   a_enc rows x H, x_r[r] padded_rows(rows) x 2H, a_r[r] padded_rows(rows) x H, a_dec rows x H.
struct FwdBufs {
    const bf16 *feat;
    bf16 *h[8];
    float *y;
    bf16 *a_enc, *a_dec;
    bf16 *x_r[4], *a_r[4];
    bool training;
};
// ev (optional, R + 3 events): recorded before the encoder, after it, after every round
// (neighbour mean and round MLP in one launch), and after the decoder.
hipError_t forward(hipStream_t st, const NetLayout &nl, const NetDev &nd, const Graph &g, uint32_t rows,
                   const FwdBufs &b, hipEvent_t *ev = nullptr);
// Loss partials (and dy) without the backward pass.
hipError_t loss_only(hipStream_t st, uint32_t rows, const float *y, const float4 *t, float inv_nfree, float *dy,
                     double *loss_part);

// Backward + weight gradients.  t: rows x 4 (target xyz, mask);  n_free: rows with mask 1.
struct BwdBufs {
    float *dy;           // rows x O   (loss gradient)
    float *dh0, *dh1;    // rows x H fp32 ping-pong
    float *dx;           // rows x 2H fp32
    bf16 *dyb;           // rows x H  bf16 gradient of a layer's output
    bf16 *dzb;           // rows x H  bf16 gradient of a layer's pre-activation
    float *partial;      // per-block weight-gradient partials
    uint32_t partial_blocks;
    float *grad;         // nparams
    double *loss_part;   // per-block loss partials (>= 1024)
};
hipError_t loss_and_backward(hipStream_t st, const NetLayout &nl, const NetDev &nd, const Graph &g, uint32_t rows,
                             const FwdBufs &f, const float4 *t, float inv_nfree, const BwdBufs &b);
uint32_t loss_blocks(uint32_t rows);

struct AdamArgs {
    uint32_t n;
    float lr, beta1, beta2, eps, bc1, bc2;
};
hipError_t launch_adam(hipStream_t st, const NetLayout &nl, const AdamArgs &a, float *theta, float *m, float *v,
                       const float *g, bf16 *w, bf16 *wt);
// Refresh the bf16 copies from theta (after the host sets parameters).
hipError_t launch_refresh(hipStream_t st, const NetLayout &nl, const float *theta, bf16 *w, bf16 *wt);

hipError_t set_lds_limits();
size_t partial_floats(const NetLayout &nl, uint32_t blocks);
uint32_t gemm_blocks(uint32_t rows);

}  // namespace ws
}  // namespace tc
