// Learned CG warm start: a per-node MLP 12 -> 128 -> 128 -> 128 -> 3 (SiLU)
// with BF16 weights and FP32 accumulation on MFMA.  It has no counterpart in
// the reference, whose CG starts from the previous velocity
// (fem_conjugate_gradients.h:48-67); the network predicts a better start and
// CG then runs to the same stopping rule, so the plain start stays the
// fallback.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

struct otfem_model;

namespace otfem {

constexpr int kMlpIn = 12;
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 3;
// FP32 master parameters, in this order: W1 [128][12], b1, W2 [128][128], b2,
// W3 [128][128], b3, W4 [3][128], b4 (row-major [out][in]).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + 128 * 12;
constexpr int kOffW2 = kOffB1 + 128;
constexpr int kOffB2 = kOffW2 + 128 * 128;
constexpr int kOffW3 = kOffB2 + 128;
constexpr int kOffB3 = kOffW3 + 128 * 128;
constexpr int kOffW4 = kOffB3 + 128;
constexpr int kOffB4 = kOffW4 + 3 * 128;
constexpr int kMlpParams = kOffB4 + 3;  // 35075

// BF16 weight fragments in MFMA operand order (8 values per lane, 64 lanes
// per fragment).  Forward: W1 (4 output tiles x 1 K-block), W2, W3 (4 x 8),
// W4 (1 x 8).  Backward (transposed): W2^T, W3^T (4 x 8), W4^T (4 x 1).
constexpr int kFragF1 = 0;
constexpr int kFragF2 = kFragF1 + 128 * 16;
constexpr int kFragF3 = kFragF2 + 128 * 128;
constexpr int kFragF4 = kFragF3 + 128 * 128;
constexpr int kFragFwd = kFragF4 + 32 * 128;  // 38912 bf16
constexpr int kFragB2 = 0;
constexpr int kFragB3 = kFragB2 + 128 * 128;
constexpr int kFragB4 = kFragB3 + 128 * 128;
constexpr int kFragBwd = kFragB4 + 128 * 16;  // 34816 bf16

// Warm-start kernels of a solver step.
struct WsArgs {
  const double *v;       // [3N] velocity v^n (internal node order)
  double *vw;            // [3N] CG start written by the prediction
  float4 *hist1;         // [N] v^{n-1} (after the step's shift: v^n)
  float4 *hist2;         // [N] v^{n-2}
  const float4 *kd;      // [N] diagonal of the rest stiffness K_ii
  const double *mass;
  const uint8_t *fixed;
  const int32_t *free_rank;  // [N] index among free nodes, -1 if fixed
  int64_t N;
  double cM, cK;
  int predict;           // run the network (else only features / history)
  int record;            // append (features, target) samples
  int64_t rec_base;      // first sample index of this step
  float *rec_X;          // [cap][16] features
  float *rec_S;          // [cap] feature scale s
  float4 *rec_T;         // [cap] target (3) and loss weight
  const void *wf;        // forward fragments
  const float *theta;    // FP32 parameters (biases are read from here)
  int *bad;              // count of nodes whose prediction was non-finite
};

hipError_t enqueue_ws_predict(const WsArgs &a, hipStream_t s);
hipError_t enqueue_ws_target(const WsArgs &a, hipStream_t s);

// What a solver needs from an attached model.
struct ModelView {
  int device;
  const void *wf;
  const float *theta;
  float *X;
  float *S;
  float4 *T;
  int64_t capacity;
  int64_t count;
};
bool model_view(otfem_model *m, ModelView *out);
// Record that `n` samples were appended by a solver step.
void model_commit(otfem_model *m, int64_t n);

// Residual operator of the residual-loss trainer: A = cM M + cK K over the
// free nodes of one mesh (caller order), block-CSR with FP32 3x3 blocks.
struct ResidualOp {
  int64_t n;                    // free nodes (samples per recorded step)
  std::vector<int64_t> row_ptr; // [n+1]
  std::vector<int32_t> cols;
  std::vector<float> blocks;    // [nnz][9] row-major
};
// Adam on the residual loss over minibatches of `steps_per_batch` recorded
// steps; returns a status and sets *err on failure.
int model_train_residual(otfem_model *m, const ResidualOp &op, int steps, int steps_per_batch,
                         double lr, uint64_t seed, float *loss, std::string *err);
// Loss and gradient of the residual loss over the given recorded steps.
int model_residual(otfem_model *m, const ResidualOp &op, int steps, int K, double lr,
                   uint64_t seed, float *loss, const int32_t *given_steps, float *grad,
                   double *loss_value, std::string *err);

}  // namespace otfem
