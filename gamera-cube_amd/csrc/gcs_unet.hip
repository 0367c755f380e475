// The ionosphere warm-start network on the device: a 4-level U-Net (32 base
// channels) whose 3x3 convolutions, their input and weight gradients run as
// implicit GEMMs on the matrix cores (bf16 operands, fp32 accumulation,
// v_mfma_f32_16x16x16_bf16), with pooling, the fused upsample+concat of the
// decoder inputs, the 1x1 head, the loss and Adam as plain kernels.
//
// Tensors are NHWC: H = colatitude rows (zero padding beyond the pole and
// the low-latitude edge), W = longitude columns (periodic), C innermost so a
// lane's four consecutive k values of an MFMA operand are one 8-byte load.
// The network is defined layer by layer in kLayers; the numpy statement of
// the same network is the oracle the tests compare with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "gcs_device.hpp"
#include "gcs_unet.hpp"

namespace gcs {
namespace unet {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kT = 256;

// float -> bf16, round to nearest even (NaN stays NaN)
__device__ __host__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline long gtid() { return blockIdx.x * (long)blockDim.x + threadIdx.x; }

// D(16x16) += A(16x16) B(16x16). Lane l supplies A[l%16][4(l/16)+0..3] and
// B[4(l/16)+0..3][l%16]; it holds D[4(l/16)+i][l%16] in element i.
__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// The input of a convolution: channels [0, c0) from src0 (at half
// resolution when up0, read with nearest upsampling), channels [c0, c0+c1)
// from src1. Both are bf16 NHWC.
struct ConvIn {
  const uint16_t *src0 = nullptr, *src1 = nullptr;
  int c0 = 0, c1 = 0, up0 = 0;
};

// Four consecutive input channels ci..ci+3 at pixel (n, yy, xx) of the
// output grid H x W; yy outside [0, H) reads zeros.
__device__ inline bf16x4 load_in4(const ConvIn &in, int n, int yy, int xx, int H, int W,
                                  int ci) {
  bf16x4 v = {0, 0, 0, 0};
  if (yy < 0 || yy >= H) return v;
  if (ci < in.c0) {
    const long p = in.up0 ? ((long)(n * (H >> 1) + (yy >> 1)) * (W >> 1) + (xx >> 1))
                          : ((long)(n * H + yy) * W + xx);
    v = *(const bf16x4 *)(in.src0 + p * in.c0 + ci);
  } else {
    const long p = (long)(n * H + yy) * W + xx;
    v = *(const bf16x4 *)(in.src1 + p * in.c1 + (ci - in.c0));
  }
  return v;
}

__device__ inline uint16_t load_in1(const ConvIn &in, int n, int yy, int xx, int H, int W,
                                    int ci) {
  if (yy < 0 || yy >= H) return 0;
  if (ci < in.c0) {
    const long p = in.up0 ? ((long)(n * (H >> 1) + (yy >> 1)) * (W >> 1) + (xx >> 1))
                          : ((long)(n * H + yy) * W + xx);
    return in.src0[p * in.c0 + ci];
  }
  const long p = (long)(n * H + yy) * W + xx;
  return in.src1[p * in.c1 + (ci - in.c0)];
}

__device__ inline int wrap(int x, int W) { return x < 0 ? x + W : (x >= W ? x - W : x); }

// 3x3 convolution as an implicit GEMM: out[m][co] = sum_k A[m][k] B[k][co]
// with m = pixel (n, y, x), k = (ky*3 + kx)*cin + ci, A = input patch,
// B[k][co] = w[co][k] (bf16 [cout][9][cin]). A block of 256 threads computes
// 64 pixels x 64 output channels; wave w owns pixels 16w..16w+15 of the
// tile. Epilogue: + bias, ReLU, bf16 store (out_bf) or fp32 store (out_f).
// Requires cin % 16 == 0, c0 % 4 == 0, cout % 16 == 0.
__global__ void __launch_bounds__(256) k_conv3(ConvIn in, int N, int H, int W,
                                               const uint16_t *w, const float *bias,
                                               int cout, int relu, uint16_t *out_bf,
                                               float *out_f) {
  const int cin = in.c0 + in.c1;
  const int M = N * H * W;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, kq = (lane >> 4) * 4;
  const int m = blockIdx.x * 64 + wave * 16 + row;
  const int mm = m < M ? m : 0;
  const int n = mm / (H * W), y = (mm / W) % H, x = mm % W;
  const int co0 = blockIdx.y * 64;
  const int nsub = min(4, (cout - co0) / 16);
  f32x4 acc[4];
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int ky = 0; ky < 3; ++ky) {
    const int yy = y + ky - 1;
    for (int kx = 0; kx < 3; ++kx) {
      const int xx = wrap(x + kx - 1, W);
      const long kbase = (long)(ky * 3 + kx) * cin;
      for (int ci0 = 0; ci0 < cin; ci0 += 16) {
        const bf16x4 a = load_in4(in, n, yy, xx, H, W, ci0 + kq);
        for (int t = 0; t < 4; ++t) {
          if (t >= nsub) break;
          const int co = co0 + 16 * t + row;
          const bf16x4 b = *(const bf16x4 *)(w + (long)co * 9 * cin + kbase + ci0 + kq);
          acc[t] = mfma16(a, b, acc[t]);
        }
      }
    }
  }
  for (int t = 0; t < nsub; ++t) {
    const int co = co0 + 16 * t + row;
    const float bco = bias != nullptr ? bias[co] : 0.f;
    for (int i = 0; i < 4; ++i) {
      const int p = blockIdx.x * 64 + wave * 16 + kq + i;
      if (p >= M) continue;
      float v = acc[t][i] + bco;
      if (relu) v = fmaxf(v, 0.f);
      if (out_bf != nullptr) out_bf[(long)p * cout + co] = to_bf16(v);
      else out_f[(long)p * cout + co] = v;
    }
  }
}

// Weight gradient: gw[co][k] += sum_m gz[m][co] A[m][k] over the pixel
// chunk [blockIdx.z * chunk, +chunk). GEMM rows = co (A operand gz^T),
// columns = k (B operand = input patch), reduction over pixels. A block of
// 256 threads owns 64 co x 64 k; wave w owns co rows 16w..16w+15. The
// block's partial sums are added to gw with fp32 atomics.
__global__ void __launch_bounds__(256) k_wgrad3(ConvIn in, int N, int H, int W,
                                                const uint16_t *gz, int cout, float *gw,
                                                int chunk) {
  const int cin = in.c0 + in.c1;
  const int K = 9 * cin;
  const int M = N * H * W;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, kq = (lane >> 4) * 4;
  const int k0 = blockIdx.x * 64;
  const int co = blockIdx.y * 64 + wave * 16 + row;  // A-operand row of this lane
  const bool co_ok = blockIdx.y * 64 + wave * 16 < cout;
  const int nsub = min(4, (K - k0) / 16);
  if (!co_ok) return;
  // this lane's B-operand columns (one per sub-tile): k -> (ky, kx, ci)
  int kky[4], kkx[4], kci[4];
  for (int t = 0; t < 4; ++t) {
    const int k = k0 + 16 * t + row;
    const int kk = k < K ? k : 0;
    const int tap = kk / cin;
    kky[t] = tap / 3;
    kkx[t] = tap % 3;
    kci[t] = kk - tap * cin;
  }
  f32x4 acc[4];
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int m_begin = blockIdx.z * chunk, m_end = min(M, m_begin + chunk);
  for (int m0 = m_begin; m0 < m_end; m0 += 16) {
    bf16x4 a;
    int pn[4], py[4], px[4];
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + kq + i;
      const bool ok = m < m_end;
      const int mm = ok ? m : m_begin;
      a[i] = ok ? (short)gz[(long)mm * cout + co] : (short)0;
      pn[i] = mm / (H * W);
      py[i] = (mm / W) % H;
      px[i] = ok ? mm % W : -1000000;  // marks a padding pixel
    }
    for (int t = 0; t < nsub; ++t) {
      bf16x4 b;
      for (int i = 0; i < 4; ++i) {
        if (px[i] < -1000) {
          b[i] = 0;
          continue;
        }
        const int yy = py[i] + kky[t] - 1, xx = wrap(px[i] + kkx[t] - 1, W);
        b[i] = (short)load_in1(in, pn[i], yy, xx, H, W, kci[t]);
      }
      acc[t] = mfma16(a, b, acc[t]);
    }
  }
  // D[co = 4(l/16)+i][k = l%16] within the wave's 16 x 64 strip
  for (int t = 0; t < nsub; ++t) {
    const int k = k0 + 16 * t + row;
    for (int i = 0; i < 4; ++i) {
      const int c = blockIdx.y * 64 + wave * 16 + kq + i;
      if (c < cout && k < K) atomicAdd(&gw[(long)c * K + k], acc[t][i]);
    }
  }
}

// Bias gradient: gb[co] += sum over the pixel chunk of gz[m][co].
__global__ void k_colsum(const uint16_t *gz, int M, int cout, float *gb, int chunk) {
  const int co = threadIdx.x;
  if (co >= cout) return;
  const int m0 = blockIdx.x * chunk, m1 = min(M, m0 + chunk);
  float s = 0.f;
  for (int m = m0; m < m1; ++m) s += from_bf16(gz[(long)m * cout + co]);
  atomicAdd(&gb[co], s);
}

// gz = bf16(g * (act > 0))
__global__ void k_relu_grad(const float *g, const uint16_t *act, long count, uint16_t *gz) {
  const long t = gtid();
  if (t >= count) return;
  gz[t] = to_bf16(from_bf16(act[t]) > 0.f ? g[t] : 0.f);
}

// 2x2 mean pool, bf16 in and out; (N, H, W) is the input grid.
__global__ void k_pool(const uint16_t *in, int N, int H, int W, int C, uint16_t *out) {
  const long t = gtid();
  const int h2 = H >> 1, w2 = W >> 1;
  if (t >= (long)N * h2 * w2 * C) return;
  const int c = (int)(t % C);
  const long p = t / C;
  const int x = (int)(p % w2), y = (int)((p / w2) % h2), n = (int)(p / ((long)w2 * h2));
  float s = 0.f;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx)
      s += from_bf16(in[(((long)n * H + 2 * y + dy) * W + 2 * x + dx) * C + c]);
  out[t] = to_bf16(0.25f * s);
}

// g_fine[n,y,x,c] += 0.25 g_coarse[n,y/2,x/2,c] (pool backward), fine grid.
__global__ void k_pool_back(const float *gc, int N, int H, int W, int C, float *gf) {
  const long t = gtid();
  if (t >= (long)N * H * W * C) return;
  const int c = (int)(t % C);
  const long p = t / C;
  const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((long)W * H));
  gf[t] += 0.25f * gc[(((long)n * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * C + c];
}

// Split the input gradient of a decoder conv (fine grid N x H x W, c0 + c1
// channels): the upsampled part sums over each 2x2 block into the coarse
// tensor's gradient, the skip part adds to the skip's gradient.
__global__ void k_split_back(const float *gcat, int N, int H, int W, int c0, int c1,
                             float *gcoarse, float *gskip) {
  const long t = gtid();
  const int h2 = H >> 1, w2 = W >> 1;
  const long ncoarse = (long)N * h2 * w2 * c0;
  const long nskip = (long)N * H * W * c1;
  const int cin = c0 + c1;
  if (t < ncoarse) {
    const int c = (int)(t % c0);
    const long p = t / c0;
    const int x = (int)(p % w2), y = (int)((p / w2) % h2), n = (int)(p / ((long)w2 * h2));
    float s = 0.f;
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx)
        s += gcat[(((long)n * H + 2 * y + dy) * W + 2 * x + dx) * cin + c];
    gcoarse[t] += s;
  } else if (t < ncoarse + nskip) {
    const long u = t - ncoarse;
    const int c = (int)(u % c1);
    const long p = u / c1;
    gskip[u] += gcat[p * cin + c0 + c];
  }
}

__global__ void k_add(const float *a, long count, float *out) {
  const long t = gtid();
  if (t < count) out[t] += a[t];
}

__global__ void k_fill(float *p, long count, float v) {
  const long t = gtid();
  if (t < count) p[t] = v;
}

// 1x1 head: out[m] = sum_c act[m][c] w[c] + b (fp32 out).
__global__ void k_head(const uint16_t *act, int M, int C, const float *w, const float *b,
                       float *out) {
  const long m = gtid();
  if (m >= M) return;
  float s = 0.f;
  for (int c = 0; c < C; ++c) s += from_bf16(act[m * C + c]) * w[c];
  out[m] = s + b[0];
}

// Head backward: g_act[m][c] = g[m] w[c]; gw[c] += sum g[m] act[m][c];
// gb += sum g[m] (block partials, then atomics).
__global__ void k_head_back(const uint16_t *act, const float *g, int M, int C, const float *w,
                            float *gact, float *gw, float *gb) {
  __shared__ float red[kT];
  const long m = gtid();
  const float gm = m < M ? g[m] : 0.f;
  if (m < M)
    for (int c = 0; c < C; ++c) gact[m * C + c] = gm * w[c];
  for (int c = 0; c <= C; ++c) {
    red[threadIdx.x] = m < M ? (c < C ? gm * from_bf16(act[m * C + c]) : gm) : 0.f;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(c < C ? &gw[c] : gb, red[0]);
    __syncthreads();
  }
}

// Mean square error: loss += sum (out - target)^2 / M; g = 2 (out - target) / M.
__global__ void k_mse(const float *out, const float *target, int M, float *g, float *loss) {
  __shared__ float red[kT];
  const long m = gtid();
  float d = 0.f;
  if (m < M) {
    d = out[m] - target[m];
    g[m] = 2.f * d / (float)M;
  }
  red[threadIdx.x] = d * d / (float)M;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Mean square of each sample's target (one block per sample).
__global__ void k_sample_energy(const float *target, int P, float *energy) {
  __shared__ float red[kT];
  const float *t = target + (long)blockIdx.x * P;
  float s = 0.f;
  for (int m = threadIdx.x; m < P; m += kT) s += t[m] * t[m];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) energy[blockIdx.x] = fmaxf(red[0] / (float)P, 1e-30f);
}

// Relative error: loss += mean over samples of mean_pixels (out - t)^2 /
// energy_b; g = 2 (out - t) / (B P energy_b).
__global__ void k_mse_rel(const float *out, const float *target, int B, int P,
                          const float *energy, float *g, float *loss) {
  __shared__ float red[kT];
  const long m = gtid();
  float d2 = 0.f;
  if (m < (long)B * P) {
    const float e = energy[m / P];
    const float d = out[m] - target[m];
    g[m] = 2.f * d / ((float)B * (float)P * e);
    d2 = d * d / ((float)B * (float)P * e);
  }
  red[threadIdx.x] = d2;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

__global__ void k_adam(float *p, const float *g, float *mo, float *ve, long count, float lr,
                       float b1, float b2, float c1, float c2, float eps) {
  const long t = gtid();
  if (t >= count) return;
  const float gt = g[t];
  const float m = b1 * mo[t] + (1.f - b1) * gt;
  const float v = b2 * ve[t] + (1.f - b2) * gt * gt;
  mo[t] = m;
  ve[t] = v;
  p[t] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// bf16 weights [cout][9][cin] and the flipped, transposed copy
// [cin][9][cout] (tap (ky, kx) -> (2 - ky, 2 - kx)) used by the input grad.
__global__ void k_prep_weights(const float *w, int cin, int cout, uint16_t *wb,
                               uint16_t *wf) {
  const long t = gtid();
  if (t >= (long)cout * 9 * cin) return;
  const int ci = (int)(t % cin);
  const int tap = (int)((t / cin) % 9);
  const int co = (int)(t / (9L * cin));
  const uint16_t v = to_bf16(w[t]);
  wb[t] = v;
  wf[((long)ci * 9 + (8 - tap)) * cout + co] = v;
}

__global__ void k_to_bf16(const float *in, long count, uint16_t *out) {
  const long t = gtid();
  if (t < count) out[t] = to_bf16(in[t]);
}

// GEMM check of the MFMA operand layout: C[M][N] = A[M][K] B[K][N], bf16 in
// fp32 out, one 16x16 tile per wave (M, N, K multiples of 16).
__global__ void k_gemm_check(const uint16_t *A, const uint16_t *B, int M, int N, int K,
                             float *C) {
  const int lane = threadIdx.x & 63;
  const int row = lane & 15, kq = (lane >> 4) * 4;
  const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
    for (int i = 0; i < 4; ++i) {
      a[i] = (short)A[(long)(m0 + row) * K + k0 + kq + i];
      b[i] = (short)B[(long)(k0 + kq + i) * N + n0 + row];
    }
    acc = mfma16(a, b, acc);
  }
  for (int i = 0; i < 4; ++i) C[(long)(m0 + kq + i) * N + n0 + row] = acc[i];
}

unsigned nb(long n, int t = kT) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ---- the network --------------------------------------------------------------

// (cin, cout, input kind, source activations); activations are numbered
// 0 = x, 1 + l = output of layer l.
const LayerSpec kLayers[kLayerCount] = {
    {"e0a", kInChannels, 32, IN_PLAIN, 0, -1},
    {"e0", 32, 32, IN_PLAIN, 1, -1},
    {"e1a", 32, 64, IN_POOL, 2, -1},
    {"e1", 64, 64, IN_PLAIN, 3, -1},
    {"e2a", 64, 128, IN_POOL, 4, -1},
    {"e2", 128, 128, IN_PLAIN, 5, -1},
    {"e3a", 128, 256, IN_POOL, 6, -1},
    {"e3", 256, 256, IN_PLAIN, 7, -1},
    {"d2a", 384, 128, IN_CAT, 8, 6},
    {"d2", 128, 128, IN_PLAIN, 9, -1},
    {"d1a", 192, 64, IN_CAT, 10, 4},
    {"d1", 64, 64, IN_PLAIN, 11, -1},
    {"d0a", 96, 32, IN_CAT, 12, 2},
    {"d0", 32, 32, IN_PLAIN, 13, -1},
};

// Level (0 = full grid) of each activation.
static int act_level(int a) {
  static const int lv[kActs] = {0, 0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0};
  return lv[a];
}
static int act_channels(int a) { return a == 0 ? kInChannels : kLayers[a - 1].cout; }

Net::~Net() {
  for (void *p : allocations) (void)hipFree(p);
}

// Device allocation recorded in net.allocations (a failed one as nullptr,
// which create() checks for).
template <typename T>
static T *dalloc(Net &net, size_t count) {
  void *p = nullptr;
  if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) p = nullptr;
  net.allocations.push_back(p);
  return (T *)p;
}

Net *create(int H, int W, int max_batch, uint64_t seed) {
  if (H <= 0 || W <= 0 || max_batch <= 0 || H % 8 != 0 || W % 8 != 0) {
    set_error("warm-start network: grid must be a positive multiple of 8 in both axes");
    return nullptr;
  }
  Net *net = new Net();
  Net &n = *net;
  n.H = H;
  n.W = W;
  n.max_batch = max_batch;
  // parameter layout: per layer w [cout][9][cin] then b [cout]; head w, b
  size_t off = 0;
  for (int l = 0; l < kLayerCount; ++l) {
    n.w_off[l] = off;
    off += (size_t)kLayers[l].cout * 9 * kLayers[l].cin;
    n.b_off[l] = off;
    off += kLayers[l].cout;
  }
  n.head_w_off = off;
  off += kHeadIn;
  n.head_b_off = off;
  off += 1;
  n.param_count = off;
  n.params = dalloc<float>(n, off);
  n.grads = dalloc<float>(n, off);
  n.adam_m = dalloc<float>(n, off);
  n.adam_v = dalloc<float>(n, off);
  size_t wcount = 0;
  for (int l = 0; l < kLayerCount; ++l) wcount += (size_t)kLayers[l].cout * 9 * kLayers[l].cin;
  n.wbf = dalloc<uint16_t>(n, wcount);
  n.wflip = dalloc<uint16_t>(n, wcount);
  size_t wo = 0;
  for (int l = 0; l < kLayerCount; ++l) {
    n.wbf_off[l] = wo;
    wo += (size_t)kLayers[l].cout * 9 * kLayers[l].cin;
  }
  long max_scratch = 0;
  for (int a = 0; a < kActs; ++a) {
    const int lv = act_level(a);
    const long pix = (long)max_batch * (H >> lv) * (W >> lv);
    n.act[a] = dalloc<uint16_t>(n, pix * act_channels(a));
    n.gact[a] = dalloc<float>(n, pix * act_channels(a));
  }
  for (int l = 0; l < kLayerCount; ++l) {
    const int lv = act_level(l + 1);
    const long pix = (long)max_batch * (H >> lv) * (W >> lv);
    max_scratch = std::max(max_scratch, pix * (long)kLayers[l].cin);
    max_scratch = std::max(max_scratch, pix * (long)kLayers[l].cout);
  }
  for (int p = 0; p < 3; ++p) {
    const long pix = (long)max_batch * (H >> (p + 1)) * (W >> (p + 1));
    n.pooled[p] = dalloc<uint16_t>(n, pix * kLayers[2 * p + 1].cout);
  }
  n.scratch = dalloc<float>(n, max_scratch);
  n.gz = dalloc<uint16_t>(n, max_scratch);
  const long pix0 = (long)max_batch * H * W;
  n.out = dalloc<float>(n, pix0);
  n.gout = dalloc<float>(n, pix0);
  n.target = dalloc<float>(n, pix0);
  n.input = dalloc<float>(n, pix0 * kInChannels);
  n.loss = dalloc<float>(n, 1);
  n.energy = dalloc<float>(n, max_batch);
  for (void *p : n.allocations)
    if (p == nullptr) {
      set_error("warm-start network: device allocation failed");
      delete net;
      return nullptr;
    }
  if (hipStreamCreateWithFlags(&n.stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("warm-start network: stream creation failed");
    delete net;
    return nullptr;
  }
  // He-normal conv weights, zero biases, a small head
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> normal(0.f, 1.f);
  std::vector<float> host(n.param_count, 0.f);
  for (int l = 0; l < kLayerCount; ++l) {
    const float sd = std::sqrt(2.f / (9.f * kLayers[l].cin));
    const size_t cnt = (size_t)kLayers[l].cout * 9 * kLayers[l].cin;
    for (size_t i = 0; i < cnt; ++i) host[n.w_off[l] + i] = normal(rng) * sd;
  }
  for (int c = 0; c < kHeadIn; ++c) host[n.head_w_off + c] = normal(rng) * 0.01f;
  if (set_params(n, host.data(), host.size()) != 0) {
    delete net;
    return nullptr;
  }
  return net;
}

void destroy(Net *net) {
  if (net == nullptr) return;
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

static int prep_weights(Net &n) {
  for (int l = 0; l < kLayerCount; ++l) {
    const LayerSpec &L = kLayers[l];
    k_prep_weights<<<nb((long)L.cout * 9 * L.cin), kT, 0, n.stream>>>(
        n.params + n.w_off[l], L.cin, L.cout, n.wbf + n.wbf_off[l], n.wflip + n.wbf_off[l]);
  }
  GCS_HIP(hipGetLastError());
  return 0;
}

int set_params(Net &n, const float *host, size_t count) {
  if (count != n.param_count) {
    set_error("warm-start network: parameter count mismatch");
    return -1;
  }
  GCS_HIP(hipMemcpyAsync(n.params, host, count * sizeof(float), hipMemcpyHostToDevice,
                         n.stream));
  k_fill<<<nb((long)count), kT, 0, n.stream>>>(n.adam_m, (long)count, 0.f);
  k_fill<<<nb((long)count), kT, 0, n.stream>>>(n.adam_v, (long)count, 0.f);
  n.adam_t = 0;
  if (prep_weights(n) != 0) return -1;
  GCS_HIP(hipStreamSynchronize(n.stream));
  return 0;
}

int get_array(Net &n, int which, float *host, size_t count) {
  if (count != n.param_count) {
    set_error("warm-start network: parameter count mismatch");
    return -1;
  }
  GCS_HIP(hipMemcpyAsync(host, which == 0 ? n.params : n.grads, count * sizeof(float),
                         hipMemcpyDeviceToHost, n.stream));
  GCS_HIP(hipStreamSynchronize(n.stream));
  return 0;
}

static ConvIn layer_input(const Net &n, int l) {
  const LayerSpec &L = kLayers[l];
  ConvIn in;
  if (L.kind == IN_PLAIN) {
    in.src1 = n.act[L.src];
    in.c1 = L.cin;
  } else if (L.kind == IN_POOL) {
    in.src1 = n.pooled[(l - 2) / 2];
    in.c1 = L.cin;
  } else {
    in.src0 = n.act[L.src];
    in.c0 = act_channels(L.src);
    in.up0 = 1;
    in.src1 = n.act[L.src2];
    in.c1 = act_channels(L.src2);
  }
  return in;
}

// Forward pass over `batch` samples whose fp32 input features are already
// in n.input ([N][H][W][kInChannels]); output in n.out.
int forward_device(Net &n, int batch) {
  hipStream_t s = n.stream;
  const long pix0 = (long)batch * n.H * n.W;
  k_to_bf16<<<nb(pix0 * kInChannels), kT, 0, s>>>(n.input, pix0 * kInChannels, n.act[0]);
  for (int l = 0; l < kLayerCount; ++l) {
    const LayerSpec &L = kLayers[l];
    const int lv = act_level(l + 1);
    const int H = n.H >> lv, W = n.W >> lv;
    if (L.kind == IN_POOL) {
      const int src = L.src;  // the pre-pool activation at level lv - 1
      k_pool<<<nb((long)batch * H * W * L.cin), kT, 0, s>>>(
          n.act[src], batch, H * 2, W * 2, L.cin, n.pooled[(l - 2) / 2]);
    }
    const ConvIn in = layer_input(n, l);
    const long M = (long)batch * H * W;
    k_conv3<<<dim3(nb(M, 64), (L.cout + 63) / 64), 256, 0, s>>>(
        in, batch, H, W, n.wbf + n.wbf_off[l], n.params + n.b_off[l], L.cout, 1,
        n.act[l + 1], nullptr);
  }
  k_head<<<nb(pix0), kT, 0, s>>>(n.act[kActs - 1], (int)pix0, kHeadIn,
                                  n.params + n.head_w_off, n.params + n.head_b_off, n.out);
  GCS_HIP(hipGetLastError());
  return 0;
}

// Backward pass from n.gout (fp32, d loss / d out); gradients in n.grads.
int backward_device(Net &n, int batch) {
  hipStream_t s = n.stream;
  const long pix0 = (long)batch * n.H * n.W;
  k_fill<<<nb((long)n.param_count), kT, 0, s>>>(n.grads, (long)n.param_count, 0.f);
  for (int a = 1; a < kActs; ++a) {
    const int lv = act_level(a);
    const long cnt = (long)batch * (n.H >> lv) * (n.W >> lv) * act_channels(a);
    k_fill<<<nb(cnt), kT, 0, s>>>(n.gact[a], cnt, 0.f);
  }
  k_head_back<<<nb(pix0), kT, 0, s>>>(n.act[kActs - 1], n.gout, (int)pix0, kHeadIn,
                                       n.params + n.head_w_off, n.gact[kActs - 1],
                                       n.grads + n.head_w_off, n.grads + n.head_b_off);
  for (int l = kLayerCount - 1; l >= 0; --l) {
    const LayerSpec &L = kLayers[l];
    const int lv = act_level(l + 1);
    const int H = n.H >> lv, W = n.W >> lv;
    const long M = (long)batch * H * W;
    k_relu_grad<<<nb(M * L.cout), kT, 0, s>>>(n.gact[l + 1], n.act[l + 1], M * L.cout, n.gz);
    // Pixel chunks small enough that the sums run in parallel: a bias sum
    // is one thread's loop over its chunk, and a weight-gradient launch has
    // (K / 64) (cout / 64) blocks per chunk, as few as five at level 0.
    const int chunk = 64;
    k_colsum<<<nb(M, chunk), 256, 0, s>>>(n.gz, (int)M, L.cout, n.grads + n.b_off[l], chunk);
    const ConvIn in = layer_input(n, l);
    const int K = 9 * L.cin;
    const int wchunk = 256;
    k_wgrad3<<<dim3((K + 63) / 64, (L.cout + 63) / 64, nb(M, wchunk)), 256, 0, s>>>(
        in, batch, H, W, n.gz, L.cout, n.grads + n.w_off[l], wchunk);
    if (l == 0) continue;  // no gradient for the input features
    // input gradient = conv of gz with the flipped, transposed weights
    ConvIn gin;
    gin.src1 = n.gz;
    gin.c1 = L.cout;
    k_conv3<<<dim3(nb(M, 64), (L.cin + 63) / 64), 256, 0, s>>>(
        gin, batch, H, W, n.wflip + n.wbf_off[l], nullptr, L.cin, 0, nullptr, n.scratch);
    if (L.kind == IN_PLAIN) {
      k_add<<<nb(M * L.cin), kT, 0, s>>>(n.scratch, M * L.cin, n.gact[L.src]);
    } else if (L.kind == IN_POOL) {
      const int src = L.src;  // fine-grid activation
      k_pool_back<<<nb(4 * M * L.cin), kT, 0, s>>>(n.scratch, batch, 2 * H, 2 * W, L.cin,
                                                    n.gact[src]);
    } else {
      const int c0 = act_channels(L.src), c1 = act_channels(L.src2);
      const long cnt = (long)batch * (H / 2) * (W / 2) * c0 + M * c1;
      k_split_back<<<nb(cnt), kT, 0, s>>>(n.scratch, batch, H, W, c0, c1, n.gact[L.src],
                                          n.gact[L.src2]);
    }
  }
  GCS_HIP(hipGetLastError());
  return 0;
}

// Loss of n.out against n.target and its gradient into n.gout.
int loss_device(Net &n, int batch) {
  const long pix0 = (long)batch * n.H * n.W;
  k_fill<<<1, 64, 0, n.stream>>>(n.loss, 1, 0.f);
  k_mse<<<nb(pix0), kT, 0, n.stream>>>(n.out, n.target, (int)pix0, n.gout, n.loss);
  GCS_HIP(hipGetLastError());
  return 0;
}

// Per-sample relative error of n.out against n.target (training loss).
int loss_relative_device(Net &n, int batch) {
  const int P = n.H * n.W;
  k_fill<<<1, 64, 0, n.stream>>>(n.loss, 1, 0.f);
  k_sample_energy<<<batch, kT, 0, n.stream>>>(n.target, P, n.energy);
  k_mse_rel<<<nb((long)batch * P), kT, 0, n.stream>>>(n.out, n.target, batch, P, n.energy,
                                                      n.gout, n.loss);
  GCS_HIP(hipGetLastError());
  return 0;
}

int adam_device(Net &n, float lr) {
  const float b1 = 0.9f, b2 = 0.999f;
  ++n.adam_t;
  const float c1 = 1.f - std::pow(b1, (float)n.adam_t);
  const float c2 = 1.f - std::pow(b2, (float)n.adam_t);
  k_adam<<<nb((long)n.param_count), kT, 0, n.stream>>>(n.params, n.grads, n.adam_m, n.adam_v,
                                                       (long)n.param_count, lr, b1, b2, c1,
                                                       c2, 1e-8f);
  return prep_weights(n);
}

int check_batch(const Net &n, int batch) {
  if (batch <= 0 || batch > n.max_batch) {
    set_error("warm-start network: batch outside [1, max_batch]");
    return -1;
  }
  return 0;
}

int gemm_check(const uint16_t *A, const uint16_t *B, int M, int N, int K, float *C) {
  if (M % 16 || N % 16 || K % 16 || M <= 0 || N <= 0 || K <= 0) {
    set_error("gemm check: sizes must be positive multiples of 16");
    return -1;
  }
  uint16_t *dA = nullptr, *dB = nullptr;
  float *dC = nullptr;
  GCS_HIP(hipMalloc(&dA, (size_t)M * K * 2));
  GCS_HIP(hipMalloc(&dB, (size_t)K * N * 2));
  GCS_HIP(hipMalloc(&dC, (size_t)M * N * 4));
  GCS_HIP(hipMemcpy(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice));
  GCS_HIP(hipMemcpy(dB, B, (size_t)K * N * 2, hipMemcpyHostToDevice));
  k_gemm_check<<<dim3(M / 16, N / 16), 64>>>(dA, dB, M, N, K, dC);
  GCS_HIP(hipGetLastError());
  GCS_HIP(hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return 0;
}

}  // namespace unet
}  // namespace gcs
