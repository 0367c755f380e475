// The ionosphere warm-start network (device side): layer table, buffers and
// the forward/backward/optimizer entry points used by the C ABI and by the
// MI solve.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "gamera_cs_unet.h"
#include "gcs_device.hpp"

namespace gcs {
namespace unet {

constexpr int kInChannels = 32;  // input features, zero-padded to 32
constexpr int kLayerCount = 14;  // 3x3 conv layers
constexpr int kActs = kLayerCount + 1;
constexpr int kHeadIn = 32;

enum InputKind { IN_PLAIN, IN_POOL, IN_CAT };

struct LayerSpec {
  const char *name;
  int cin, cout;
  int kind;
  int src, src2;  // activation indices (0 = input, 1 + l = layer l output)
};

extern const LayerSpec kLayers[kLayerCount];

struct Net {
  int H = 0, W = 0, max_batch = 0;
  size_t param_count = 0;
  size_t w_off[kLayerCount] = {}, b_off[kLayerCount] = {}, wbf_off[kLayerCount] = {};
  size_t head_w_off = 0, head_b_off = 0;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  int adam_t = 0;
  uint16_t *wbf = nullptr, *wflip = nullptr;  // bf16 weights, flipped-transposed copy
  uint16_t *act[kActs] = {};                  // bf16 activations
  float *gact[kActs] = {};                    // fp32 activation gradients
  uint16_t *pooled[3] = {};                   // pooled encoder outputs
  float *scratch = nullptr;                   // input gradient of one layer
  uint16_t *gz = nullptr;                     // masked output gradient, bf16
  float *input = nullptr;                     // fp32 features [N][H][W][16]
  float *out = nullptr, *gout = nullptr, *target = nullptr;
  float *loss = nullptr;
  float *energy = nullptr;                    // per-sample target mean square
  // potential = out * out_scale * R^2 * scale / SigmaP (set on the first
  // training batch so that targets have unit RMS)
  double out_scale = 1.0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocations;
  ~Net();
};

Net *create(int H, int W, int max_batch, uint64_t seed);
void destroy(Net *net);
int set_params(Net &n, const float *host, size_t count);
int get_array(Net &n, int which, float *host, size_t count);  // 0 params, 1 grads
int check_batch(const Net &n, int batch);
int forward_device(Net &n, int batch);
int loss_device(Net &n, int batch);
int loss_relative_device(Net &n, int batch);
int backward_device(Net &n, int batch);
int adam_device(Net &n, float lr);
int gemm_check(const uint16_t *A, const uint16_t *B, int M, int N, int K, float *C);

}  // namespace unet

// MI side (gcs_mi.hip): attach the network, generate pairs, train, and one
// solve from host arrays.
int mi_attach_warm_start(Problem &P, unet::Net *net);
int mi_make_pairs(Problem &P, uint64_t seed, int count, double sp, double sh, double *fac_out,
                  double *phi_out);
int mi_train_warm_start(Problem &P, unet::Net &n, const gcs_warm_train_config &t,
                        float *loss_history);
int mi_train_warm_start_on(Problem &P, unet::Net &n, const gcs_warm_train_config &t,
                           const double *fac, const double *phi, const int *hemi, int count,
                           float *loss_history);
int mi_solve_current(Problem &P, const double *fac, double *potential, int use_net,
                     int iterations[2], double *ms);

}  // namespace gcs
