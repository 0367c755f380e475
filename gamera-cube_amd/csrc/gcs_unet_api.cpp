// C ABI of the warm-start network (include/gamera_cs_unet.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "gamera_cs_unet.h"
#include "gcs_unet.hpp"

using gcs::Problem;
using gcs::unet::Net;

namespace {
Net *as_net(gcs_unet *p) { return reinterpret_cast<Net *>(p); }
const Net *as_net(const gcs_unet *p) { return reinterpret_cast<const Net *>(p); }
Problem *as_problem(gcs_problem *p) { return reinterpret_cast<Problem *>(p); }

int upload(Net &n, const float *host, float *dev, size_t count) {
  if (hipMemcpyAsync(dev, host, count * sizeof(float), hipMemcpyHostToDevice, n.stream) !=
      hipSuccess) {
    gcs::set_error("warm-start network: upload failed");
    return -1;
  }
  return 0;
}

int download(Net &n, const float *dev, float *host, size_t count) {
  if (hipMemcpyAsync(host, dev, count * sizeof(float), hipMemcpyDeviceToHost, n.stream) !=
          hipSuccess ||
      hipStreamSynchronize(n.stream) != hipSuccess) {
    gcs::set_error("warm-start network: download failed");
    return -1;
  }
  return 0;
}
}  // namespace

extern "C" {

gcs_unet *gcs_unet_create(int colatitude_points, int azimuth_points, int max_batch,
                          unsigned long long seed) {
  return reinterpret_cast<gcs_unet *>(
      gcs::unet::create(colatitude_points, azimuth_points, max_batch, seed));
}

void gcs_unet_destroy(gcs_unet *net) { gcs::unet::destroy(as_net(net)); }

size_t gcs_unet_param_count(const gcs_unet *net) {
  return net == nullptr ? 0 : as_net(net)->param_count;
}

int gcs_unet_set_params(gcs_unet *net, const float *params, size_t count) {
  if (net == nullptr || params == nullptr) return -1;
  return gcs::unet::set_params(*as_net(net), params, count);
}

int gcs_unet_get_params(gcs_unet *net, float *params, size_t count) {
  if (net == nullptr || params == nullptr) return -1;
  return gcs::unet::get_array(*as_net(net), 0, params, count);
}

int gcs_unet_get_grads(gcs_unet *net, float *grads, size_t count) {
  if (net == nullptr || grads == nullptr) return -1;
  return gcs::unet::get_array(*as_net(net), 1, grads, count);
}

double gcs_unet_output_scale(const gcs_unet *net) {
  return net == nullptr ? 0.0 : as_net(net)->out_scale;
}

int gcs_unet_set_output_scale(gcs_unet *net, double scale) {
  if (net == nullptr || !(scale > 0.0)) return -1;
  as_net(net)->out_scale = scale;
  return 0;
}

int gcs_unet_forward(gcs_unet *net, const float *features, int batch, float *out) {
  if (net == nullptr || features == nullptr || out == nullptr) return -1;
  Net &n = *as_net(net);
  if (gcs::unet::check_batch(n, batch) != 0) return -1;
  const size_t pix = (size_t)batch * n.H * n.W;
  if (upload(n, features, n.input, pix * gcs::unet::kInChannels) != 0 ||
      gcs::unet::forward_device(n, batch) != 0)
    return -1;
  return download(n, n.out, out, pix);
}

int gcs_unet_gradients(gcs_unet *net, const float *features, const float *target, int batch,
                       double *loss) {
  if (net == nullptr || features == nullptr || target == nullptr) return -1;
  Net &n = *as_net(net);
  if (gcs::unet::check_batch(n, batch) != 0) return -1;
  const size_t pix = (size_t)batch * n.H * n.W;
  if (upload(n, features, n.input, pix * gcs::unet::kInChannels) != 0 ||
      upload(n, target, n.target, pix) != 0 || gcs::unet::forward_device(n, batch) != 0 ||
      gcs::unet::loss_device(n, batch) != 0 || gcs::unet::backward_device(n, batch) != 0)
    return -1;
  float l = 0.f;
  if (download(n, n.loss, &l, 1) != 0) return -1;
  if (loss != nullptr) *loss = l;
  return 0;
}

int gcs_unet_adam_step(gcs_unet *net, double learning_rate) {
  if (net == nullptr || !(learning_rate > 0.0)) return -1;
  Net &n = *as_net(net);
  if (gcs::unet::adam_device(n, (float)learning_rate) != 0) return -1;
  if (hipStreamSynchronize(n.stream) != hipSuccess) {
    gcs::set_error("warm-start network: stream error");
    return -1;
  }
  return 0;
}

int gcs_mfma_gemm_check(const uint16_t *a, const uint16_t *b, int m, int n, int k, float *c) {
  if (a == nullptr || b == nullptr || c == nullptr) return -1;
  return gcs::unet::gemm_check(a, b, m, n, k, c);
}

int gcs_problem_mi_attach_warm_start(gcs_problem *problem, gcs_unet *net) {
  if (problem == nullptr) return -1;
  return gcs::mi_attach_warm_start(*as_problem(problem), as_net(net));
}

int gcs_problem_mi_train_warm_start(gcs_problem *problem, gcs_unet *net,
                                    const gcs_warm_train_config *cfg, float *loss_history) {
  if (problem == nullptr || net == nullptr || cfg == nullptr) return -1;
  return gcs::mi_train_warm_start(*as_problem(problem), *as_net(net), *cfg, loss_history);
}

int gcs_problem_mi_train_warm_start_on(gcs_problem *problem, gcs_unet *net,
                                       const gcs_warm_train_config *cfg, const double *current,
                                       const double *potential, const int *hemisphere,
                                       int count, float *loss_history) {
  if (problem == nullptr || net == nullptr || cfg == nullptr) return -1;
  return gcs::mi_train_warm_start_on(*as_problem(problem), *as_net(net), *cfg, current,
                                     potential, hemisphere, count, loss_history);
}

int gcs_problem_mi_make_pairs(gcs_problem *problem, unsigned long long seed, int count,
                              double pedersen, double hall, double *current,
                              double *potential) {
  if (problem == nullptr || current == nullptr || potential == nullptr) return -1;
  return gcs::mi_make_pairs(*as_problem(problem), seed, count, pedersen, hall, current,
                            potential);
}

int gcs_problem_mi_solve_current(gcs_problem *problem, const double *current,
                                 double *potential, int use_network, int iterations[2],
                                 double *milliseconds) {
  if (problem == nullptr || current == nullptr || potential == nullptr ||
      iterations == nullptr || milliseconds == nullptr)
    return -1;
  return gcs::mi_solve_current(*as_problem(problem), current, potential, use_network,
                               iterations, milliseconds);
}

}  // extern "C"
