// Learned warm start for the CDMFT bath fit: a small MLP n_in -> H -> H -> n_out
// (tanh hidden layers, linear output), trained and evaluated on the device.
//
//  * FP32 forward and backward in one fused kernel per tile of samples: the weights of
//    the wide layer sit in LDS, the activations of the tile stay in LDS for the backward
//    pass; each tile writes its partial weight gradient, and a second kernel sums the
//    partials in tile order (deterministic).
//  * Adam update as one elementwise kernel.
//  * BF16 inference on the matrix cores: both hidden-to-hidden and hidden-to-output
//    products run as v_mfma_f32_16x16x32_bf16 on 16-sample tiles (FP32 accumulation,
//    FP32 bias and tanh); the first layer (n_in inputs) stays FP32.
//
// Parameter layout (float): W1 [H][n_in], b1 [H], W2 [H][H], b2 [H], W3 [n_out][H], b3.
// Loss: mean over samples and outputs of (y - t)^2.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "../../include/qed.h"
#include "qed_common.hpp"

struct qed_mlp {
  int n_in = 0, n_h = 0, n_out = 0;
  size_t n_params = 0;
  qed::DevBuf params, m, v, grad;  // fp32
  qed::DevBuf wbf;                 // bf16 copies of W2, W3 (inference)
  bool wbf_valid = false;
  long step = 0;
};

namespace qed {
namespace {

constexpr int kTile = 16;      // samples per workgroup tile
constexpr int kThreads = 256;

struct Off {
  size_t W1, b1, W2, b2, W3, b3, total;
};

__host__ __device__ Off offsets(int ni, int nh, int no) {
  Off o;
  o.W1 = 0;
  o.b1 = o.W1 + (size_t)nh * ni;
  o.W2 = o.b1 + nh;
  o.b2 = o.W2 + (size_t)nh * nh;
  o.W3 = o.b2 + nh;
  o.b3 = o.W3 + (size_t)no * nh;
  o.total = o.b3 + no;
  return o;
}

struct MlpView {
  int ni, nh, no;
  const float *W1, *b1, *W2, *b2, *W3, *b3;
};

MlpView view(const qed_mlp& m) {
  Off o = offsets(m.n_in, m.n_h, m.n_out);
  const float* p = m.params.as<float>();
  return {m.n_in, m.n_h, m.n_out, p + o.W1, p + o.b1, p + o.W2, p + o.b2, p + o.W3, p + o.b3};
}

// forward of one tile into LDS: x (ns x ni), h1, h2 (ns x nh), y (ns x no)
__device__ void tile_forward(const MlpView& V, const float* W2s, const float* xs, float* h1,
                             float* h2, float* ys, int ns) {
  const int t = threadIdx.x;
  for (int q = t; q < ns * V.nh; q += blockDim.x) {
    const int s = q / V.nh, j = q % V.nh;
    float a = V.b1[j];
    for (int k = 0; k < V.ni; ++k) a += V.W1[j * V.ni + k] * xs[s * V.ni + k];
    h1[q] = tanhf(a);
  }
  __syncthreads();
  for (int q = t; q < ns * V.nh; q += blockDim.x) {
    const int s = q / V.nh, j = q % V.nh;
    float a = V.b2[j];
    const float* w = W2s + j * V.nh;
    const float* h = h1 + s * V.nh;
    for (int k = 0; k < V.nh; ++k) a += w[k] * h[k];
    h2[q] = tanhf(a);
  }
  __syncthreads();
  for (int q = t; q < ns * V.no; q += blockDim.x) {
    const int s = q / V.no, o = q % V.no;
    float a = V.b3[o];
    const float* w = V.W3 + o * V.nh;
    const float* h = h2 + s * V.nh;
    for (int k = 0; k < V.nh; ++k) a += w[k] * h[k];
    ys[q] = a;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_forward(MlpView V, int n, const float* __restrict__ x,
                                                      float* __restrict__ y) {
  extern __shared__ float sm[];
  float* W2s = sm;
  float* xs = W2s + V.nh * V.nh;
  float* h1 = xs + kTile * V.ni;
  float* h2 = h1 + kTile * V.nh;
  float* ys = h2 + kTile * V.nh;
  for (int i = threadIdx.x; i < V.nh * V.nh; i += blockDim.x) W2s[i] = V.W2[i];
  const int s0 = blockIdx.x * kTile;
  const int ns = min(kTile, n - s0);
  for (int i = threadIdx.x; i < ns * V.ni; i += blockDim.x) xs[i] = x[(size_t)s0 * V.ni + i];
  __syncthreads();
  tile_forward(V, W2s, xs, h1, h2, ys, ns);
  for (int i = threadIdx.x; i < ns * V.no; i += blockDim.x) y[(size_t)s0 * V.no + i] = ys[i];
}

// forward + backward of one tile; partial gradient of the tile -> part[tile][param],
// partial loss -> lpart[tile]. inv = 1/(n * no) (mean over the whole batch).
__global__ __launch_bounds__(kThreads) void k_loss_grad(MlpView V, int n, const float* __restrict__ x,
                                                        const float* __restrict__ tgt, float inv,
                                                        size_t np, float* __restrict__ part,
                                                        float* __restrict__ lpart) {
  extern __shared__ float sm[];
  float* W2s = sm;
  float* xs = W2s + V.nh * V.nh;
  float* h1 = xs + kTile * V.ni;
  float* h2 = h1 + kTile * V.nh;
  float* ys = h2 + kTile * V.nh;
  float* dz2 = ys + kTile * V.no;
  float* dz1 = dz2 + kTile * V.nh;
  float* red = dz1 + kTile * V.nh;  // kThreads floats
  for (int i = threadIdx.x; i < V.nh * V.nh; i += blockDim.x) W2s[i] = V.W2[i];
  const int tile = blockIdx.x;
  const int s0 = tile * kTile;
  const int ns = min(kTile, n - s0);
  for (int i = threadIdx.x; i < ns * V.ni; i += blockDim.x) xs[i] = x[(size_t)s0 * V.ni + i];
  __syncthreads();
  tile_forward(V, W2s, xs, h1, h2, ys, ns);
  const int t = threadIdx.x;
  // dy = 2 (y - t) inv (kept in ys), loss partial
  float lsum = 0.0f;
  for (int q = t; q < ns * V.no; q += blockDim.x) {
    const float d = ys[q] - tgt[(size_t)s0 * V.no + q];
    lsum += d * d;
    ys[q] = 2.0f * d * inv;
  }
  red[t] = lsum;
  __syncthreads();
  if (t == 0) {
    float s = 0.0f;
    for (int i = 0; i < blockDim.x; ++i) s += red[i];
    lpart[tile] = s * inv;
  }
  // dz2 = (W3^T dy) * (1 - h2^2)
  for (int q = t; q < ns * V.nh; q += blockDim.x) {
    const int s = q / V.nh, j = q % V.nh;
    float a = 0.0f;
    for (int o = 0; o < V.no; ++o) a += V.W3[o * V.nh + j] * ys[s * V.no + o];
    dz2[q] = a * (1.0f - h2[q] * h2[q]);
  }
  __syncthreads();
  // dz1 = (W2^T dz2) * (1 - h1^2)
  for (int q = t; q < ns * V.nh; q += blockDim.x) {
    const int s = q / V.nh, k = q % V.nh;
    float a = 0.0f;
    for (int j = 0; j < V.nh; ++j) a += W2s[j * V.nh + k] * dz2[s * V.nh + j];
    dz1[q] = a * (1.0f - h1[q] * h1[q]);
  }
  __syncthreads();
  const Off o = offsets(V.ni, V.nh, V.no);
  float* g = part + (size_t)tile * np;
  for (size_t q = t; q < (size_t)V.nh * V.ni; q += blockDim.x) {
    const int j = (int)(q / V.ni), k = (int)(q % V.ni);
    float a = 0.0f;
    for (int s = 0; s < ns; ++s) a += dz1[s * V.nh + j] * xs[s * V.ni + k];
    g[o.W1 + q] = a;
  }
  for (int j = t; j < V.nh; j += blockDim.x) {
    float a = 0.0f, b = 0.0f;
    for (int s = 0; s < ns; ++s) {
      a += dz1[s * V.nh + j];
      b += dz2[s * V.nh + j];
    }
    g[o.b1 + j] = a;
    g[o.b2 + j] = b;
  }
  for (size_t q = t; q < (size_t)V.nh * V.nh; q += blockDim.x) {
    const int j = (int)(q / V.nh), k = (int)(q % V.nh);
    float a = 0.0f;
    for (int s = 0; s < ns; ++s) a += dz2[s * V.nh + j] * h1[s * V.nh + k];
    g[o.W2 + q] = a;
  }
  for (size_t q = t; q < (size_t)V.no * V.nh; q += blockDim.x) {
    const int oo = (int)(q / V.nh), j = (int)(q % V.nh);
    float a = 0.0f;
    for (int s = 0; s < ns; ++s) a += ys[s * V.no + oo] * h2[s * V.nh + j];
    g[o.W3 + q] = a;
  }
  for (int oo = t; oo < V.no; oo += blockDim.x) {
    float a = 0.0f;
    for (int s = 0; s < ns; ++s) a += ys[s * V.no + oo];
    g[o.b3 + oo] = a;
  }
}

// grad[p] = sum over tiles of part[tile][p] (tile order); loss = sum of lpart
__global__ void k_reduce_tiles(int ntiles, size_t np, const float* __restrict__ part,
                               const float* __restrict__ lpart, float* __restrict__ grad,
                               float* __restrict__ loss) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) {
    float a = 0.0f;
    for (int i = 0; i < ntiles; ++i) a += part[(size_t)i * np + p];
    grad[p] = a;
  }
  if (p == 0) {
    float l = 0.0f;
    for (int i = 0; i < ntiles; ++i) l += lpart[i];
    *loss = l;
  }
}

__global__ void k_adam(size_t np, float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float c1, float c2,
                       float eps) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1.0f - b1) * gi;
  const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  w[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- BF16 inference on the matrix cores ------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  if ((u & 0x7f800000u) == 0x7f800000u) r = u;  // inf / nan unchanged
  uint16_t h = (uint16_t)(r >> 16);
  __bf16 b;
  __builtin_memcpy(&b, &h, 2);
  return b;
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ src, __bf16* __restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = to_bf16(src[i]);
}

// One wavefront per 16-sample tile. Hidden width H (multiple of 32) and output width
// n_out (multiple of 16) go through v_mfma_f32_16x16x32_bf16:
//   A (16 x 32): lane l holds row l%16, k = 8 (l/16) .. +7
//   B (32 x 16): lane l holds column l%16, k = 8 (l/16) .. +7
//   D (16 x 16): lane l holds column l%16, rows 4 (l/16) .. +3
// Activations are kept in LDS as bf16 (row-major, 16 x H).
__global__ __launch_bounds__(64) void k_forward_bf16(MlpView V, const __bf16* __restrict__ W2b,
                                                     const __bf16* __restrict__ W3b, int n,
                                                     const float* __restrict__ x,
                                                     float* __restrict__ y) {
  extern __shared__ __bf16 hs[];  // h1 (16 x H) then h2 (16 x H)
  __bf16* h1 = hs;
  __bf16* h2 = hs + 16 * V.nh;
  const int lane = threadIdx.x;
  const int s0 = blockIdx.x * 16;
  // layer 1 in FP32 (n_in inputs)
  for (int q = lane; q < 16 * V.nh; q += 64) {
    const int s = q / V.nh, j = q % V.nh;
    float a = 0.0f;
    if (s0 + s < n) {
      a = V.b1[j];
      for (int k = 0; k < V.ni; ++k) a += V.W1[j * V.ni + k] * x[(size_t)(s0 + s) * V.ni + k];
      a = tanhf(a);
    }
    h1[q] = to_bf16(a);
  }
  __syncthreads();
  const int row = lane % 16, kq = (lane / 16) * 8;
  // layer 2: h2[s][j] = tanh(sum_k W2[j][k] h1[s][k] + b2[j]); B[k][j] = W2[j][k]
  for (int jt = 0; jt < V.nh; jt += 16) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < V.nh; k0 += 32) {
      bf16x8 a, b;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        a[i] = h1[row * V.nh + k0 + kq + i];
        b[i] = W2b[(size_t)(jt + row) * V.nh + k0 + kq + i];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int s = (lane / 16) * 4 + i, j = jt + row;
      h2[s * V.nh + j] = to_bf16(tanhf(acc[i] + V.b2[j]));
    }
  }
  __syncthreads();
  // layer 3: y[s][o] = sum_k W3[o][k] h2[s][k] + b3[o]
  for (int ot = 0; ot < V.no; ot += 16) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < V.nh; k0 += 32) {
      bf16x8 a, b;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        a[i] = h2[row * V.nh + k0 + kq + i];
        b[i] = W3b[(size_t)(ot + row) * V.nh + k0 + kq + i];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int s = (lane / 16) * 4 + i, o = ot + row;
      if (s0 + s < n) y[(size_t)(s0 + s) * V.no + o] = acc[i] + V.b3[o];
    }
  }
}

size_t lds_fp32(const qed_mlp& m, bool backward) {
  size_t f = (size_t)m.n_h * m.n_h + kTile * (m.n_in + 2 * m.n_h + m.n_out);
  if (backward) f += 2 * (size_t)kTile * m.n_h + kThreads;
  return f * sizeof(float);
}

void check_lds(size_t bytes) {
  if (bytes > 160 * 1024) fail("MLP too wide for one workgroup's LDS");
}

template <class T>
DevBuf upload(const T* p, size_t n) {
  DevBuf b(std::max<size_t>(n, 1) * sizeof(T));
  if (n) QED_HIP(hipMemcpyAsync(b.p, p, n * sizeof(T), hipMemcpyHostToDevice, stream()));
  return b;
}

void set_attr(const void* fn, size_t bytes) {
  QED_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// loss and gradient on device buffers (x, t already on the device); grad -> m.grad
float loss_grad_dev(qed_mlp& m, int n, const float* dx, const float* dt, DevBuf& part,
                    DevBuf& lpart, DevBuf& dloss, bool read_loss) {
  const int ntiles = (n + kTile - 1) / kTile;
  const size_t np = m.n_params;
  if (part.bytes < (size_t)ntiles * np * sizeof(float)) part.alloc((size_t)ntiles * np * sizeof(float));
  if (lpart.bytes < (size_t)ntiles * sizeof(float)) lpart.alloc((size_t)ntiles * sizeof(float));
  if (!dloss.p) dloss.alloc(sizeof(float));
  const size_t lds = lds_fp32(m, true);
  set_attr((const void*)k_loss_grad, lds);
  k_loss_grad<<<ntiles, kThreads, lds, stream()>>>(view(m), n, dx, dt, 1.0f / ((float)n * m.n_out),
                                                   np, part.as<float>(), lpart.as<float>());
  k_reduce_tiles<<<(unsigned)((np + 255) / 256), 256, 0, stream()>>>(
      ntiles, np, part.as<float>(), lpart.as<float>(), m.grad.as<float>(), dloss.as<float>());
  QED_HIP(hipGetLastError());
  float loss = 0.0f;
  if (read_loss) {
    QED_HIP(hipMemcpyAsync(&loss, dloss.p, sizeof(float), hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  }
  return loss;
}

void check_model(const qed_mlp* m) {
  if (!m) fail("MLP handle is null");
}

}  // namespace
}  // namespace qed

using namespace qed;

extern "C" {

int qed_mlp_create(int n_in, int n_hidden, int n_out, uint64_t seed, qed_mlp** out) {
  return guard([&] {
    if (!out) fail("mlp_create: null output");
    if (n_in <= 0 || n_hidden <= 0 || n_out <= 0) fail("mlp_create: sizes must be positive");
    std::unique_ptr<qed_mlp> m(new qed_mlp);
    m->n_in = n_in;
    m->n_h = n_hidden;
    m->n_out = n_out;
    const Off o = offsets(n_in, n_hidden, n_out);
    m->n_params = o.total;
    check_lds(lds_fp32(*m, true));
    // Glorot-uniform weights, zero biases, from a seeded generator
    std::vector<float> p(o.total, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, int fan_out, int fan_in) {
      const float lim = std::sqrt(6.0f / (float)(fan_in + fan_out));
      std::uniform_real_distribution<float> u(-lim, lim);
      for (size_t i = 0; i < (size_t)fan_out * fan_in; ++i) p[off + i] = u(rng);
    };
    fill(o.W1, n_hidden, n_in);
    fill(o.W2, n_hidden, n_hidden);
    fill(o.W3, n_out, n_hidden);
    m->params = upload(p.data(), p.size());
    m->m.alloc(o.total * sizeof(float));
    m->v.alloc(o.total * sizeof(float));
    m->grad.alloc(o.total * sizeof(float));
    QED_HIP(hipMemsetAsync(m->m.p, 0, o.total * sizeof(float), stream()));
    QED_HIP(hipMemsetAsync(m->v.p, 0, o.total * sizeof(float), stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    *out = m.release();
  });
}

int qed_mlp_destroy(qed_mlp* m) {
  return guard([&] {
    QED_HIP(hipStreamSynchronize(stream()));
    delete m;
  });
}

int qed_mlp_get_params(qed_mlp* m, float* params, int* count) {
  return guard([&] {
    check_model(m);
    if (count) *count = (int)m->n_params;
    if (params) {
      QED_HIP(hipMemcpyAsync(params, m->params.p, m->n_params * sizeof(float), hipMemcpyDeviceToHost,
                             stream()));
      QED_HIP(hipStreamSynchronize(stream()));
    }
  });
}

int qed_mlp_set_params(qed_mlp* m, const float* params, int count) {
  return guard([&] {
    check_model(m);
    if (count != (int)m->n_params) fail("mlp_set_params: wrong parameter count");
    QED_HIP(hipMemcpyAsync(m->params.p, params, m->n_params * sizeof(float), hipMemcpyHostToDevice,
                           stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    m->wbf_valid = false;
  });
}

int qed_mlp_forward(qed_mlp* m, int n, const float* x, float* y) {
  return guard([&] {
    check_model(m);
    if (n <= 0) return;
    DevBuf dx = upload(x, (size_t)n * m->n_in), dy((size_t)n * m->n_out * sizeof(float));
    set_attr((const void*)k_forward, lds_fp32(*m, false));
    k_forward<<<(n + kTile - 1) / kTile, kThreads, lds_fp32(*m, false), stream()>>>(
        view(*m), n, dx.as<float>(), dy.as<float>());
    QED_HIP(hipGetLastError());
    QED_HIP(hipMemcpyAsync(y, dy.p, (size_t)n * m->n_out * sizeof(float), hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_mlp_forward_bf16(qed_mlp* m, int n, const float* x, float* y) {
  return guard([&] {
    check_model(m);
    if (n <= 0) return;
    if (m->n_h % 32 || m->n_out % 16)
      fail("mlp_forward_bf16: hidden width must be a multiple of 32 and outputs of 16");
    const Off o = offsets(m->n_in, m->n_h, m->n_out);
    const size_t nw2 = (size_t)m->n_h * m->n_h, nw3 = (size_t)m->n_out * m->n_h;
    if (!m->wbf_valid) {
      if (m->wbf.bytes < (nw2 + nw3) * 2) m->wbf.alloc((nw2 + nw3) * 2);
      const float* p = m->params.as<float>();
      __bf16* w = m->wbf.as<__bf16>();
      k_to_bf16<<<(unsigned)((nw2 + 255) / 256), 256, 0, stream()>>>(nw2, p + o.W2, w);
      k_to_bf16<<<(unsigned)((nw3 + 255) / 256), 256, 0, stream()>>>(nw3, p + o.W3, w + nw2);
      m->wbf_valid = true;
    }
    DevBuf dx = upload(x, (size_t)n * m->n_in), dy((size_t)n * m->n_out * sizeof(float));
    const __bf16* w = m->wbf.as<__bf16>();
    k_forward_bf16<<<(n + 15) / 16, 64, 2 * 16 * m->n_h * sizeof(__bf16), stream()>>>(
        view(*m), w, w + nw2, n, dx.as<float>(), dy.as<float>());
    QED_HIP(hipGetLastError());
    QED_HIP(hipMemcpyAsync(y, dy.p, (size_t)n * m->n_out * sizeof(float), hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_mlp_loss_grad(qed_mlp* m, int n, const float* x, const float* t, float* loss, float* grad) {
  return guard([&] {
    check_model(m);
    if (n <= 0) fail("mlp_loss_grad: empty batch");
    DevBuf dx = upload(x, (size_t)n * m->n_in), dt = upload(t, (size_t)n * m->n_out);
    DevBuf part, lpart, dloss;
    float l = loss_grad_dev(*m, n, dx.as<float>(), dt.as<float>(), part, lpart, dloss, true);
    if (loss) *loss = l;
    if (grad) {
      QED_HIP(hipMemcpyAsync(grad, m->grad.p, m->n_params * sizeof(float), hipMemcpyDeviceToHost, stream()));
      QED_HIP(hipStreamSynchronize(stream()));
    }
  });
}

int qed_mlp_train(qed_mlp* m, int n, const float* x, const float* t, int epochs, int bs, float lr,
                  uint64_t seed, float* loss_hist) {
  return guard([&] {
    check_model(m);
    if (n <= 0 || epochs <= 0 || bs <= 0) fail("mlp_train: bad sizes");
    // the data set stays resident; minibatches are gathered by a host-side permutation
    // into contiguous staging buffers
    const int ni = m->n_in, no = m->n_out;
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::mt19937_64 rng(seed);
    std::vector<float> xb((size_t)bs * ni), tb((size_t)bs * no);
    DevBuf dxb((size_t)bs * ni * 4), dtb((size_t)bs * no * 4), part, lpart, dloss;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<float> hist(epochs, 0.0f);
    for (int ep = 0; ep < epochs; ++ep) {
      std::shuffle(perm.begin(), perm.end(), rng);
      double acc = 0.0;
      int nb = 0;
      for (int b0 = 0; b0 < n; b0 += bs) {
        const int cnt = std::min(bs, n - b0);
        for (int i = 0; i < cnt; ++i) {
          std::memcpy(&xb[(size_t)i * ni], x + (size_t)perm[b0 + i] * ni, ni * 4);
          std::memcpy(&tb[(size_t)i * no], t + (size_t)perm[b0 + i] * no, no * 4);
        }
        QED_HIP(hipMemcpyAsync(dxb.p, xb.data(), (size_t)cnt * ni * 4, hipMemcpyHostToDevice, stream()));
        QED_HIP(hipMemcpyAsync(dtb.p, tb.data(), (size_t)cnt * no * 4, hipMemcpyHostToDevice, stream()));
        float l = loss_grad_dev(*m, cnt, dxb.as<float>(), dtb.as<float>(), part, lpart, dloss, true);
        acc += l;
        ++nb;
        ++m->step;
        const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
        k_adam<<<(unsigned)((m->n_params + 255) / 256), 256, 0, stream()>>>(
            m->n_params, m->params.as<float>(), m->m.as<float>(), m->v.as<float>(), m->grad.as<float>(),
            lr, b1, b2, c1, c2, eps);
        QED_HIP(hipGetLastError());
      }
      hist[ep] = (float)(acc / nb);
    }
    QED_HIP(hipStreamSynchronize(stream()));
    m->wbf_valid = false;
    if (loss_hist) std::memcpy(loss_hist, hist.data(), epochs * sizeof(float));
  });
}

}  // extern "C"
