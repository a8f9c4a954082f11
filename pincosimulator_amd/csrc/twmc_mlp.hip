// twmc_mlp.hip -- the warm start's network: a small fully connected MLP
// (inputs -> hidden layers with exact-erf GELU -> linear outputs) trained and
// evaluated by the kernels in this file, FP32 throughout, behind the C ABI
// twmc_mlp_* of include/twmc_hip.h.  No reference counterpart: the network
// maps drive / lattice parameters to a Gaussian steady-state estimate (mean
// field + radial k-space populations) that seeds TWMC_INIT_GAUSSIAN.
//
// Layout: activations row-major [rows][features]; layer l holds W_l
// [in_l][out_l] row-major followed by b_l [out_l] in one flat FP32 parameter
// vector, so Z_l = A_{l-1} W_l + b_l is a plain NN product.
//
// Kernels:
//   mlp_gemm_kernel<MODE>  C[M x N] = op(A)[M x K] op(B)[K x N] with strided
//                          operands (covers NN forward, NT for dA = dZ W^T, TN
//                          for dW = A^T dZ), 64x64 tile per 256-thread
//                          workgroup, K staged through LDS 16 at a time, 4x4
//                          outputs per thread; epilogues: plain, +bias,
//                          +bias then GELU (keeping the pre-activation), and
//                          times GELU'(Z) for the backward pass
//   mlp_colsum_kernel      bias gradient, fixed summation order
//   mlp_loss_kernel        weighted MSE and its gradient, one workgroup,
//                          fixed-order FP64 reduction
//   mlp_adam_kernel        Adam update with bias correction
//   mlp_gather_kernel      minibatch rows by index
// Every reduction runs in a fixed order, so training is bitwise repeatable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/twmc_hip.h"
#include "twmc_rng.h"

namespace twmc {
namespace {

constexpr int kTile = 64;
constexpr int kTileK = 16;
constexpr int kMaxLayers = 16;

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_BACK = 3 };

struct GemmArgs {
  int M, N, K;
  const float* A;
  long sam, sak;  // A(m, k) = A[m sam + k sak]
  const float* B;
  long sbk, sbn;  // B(k, n) = B[k sbk + n sbn]
  float* C;       // row-major M x N
  const float* bias;  // [N]
  float* Z;           // EPI_BIAS_GELU: pre-activation out, row-major M x N
  const float* Zin;   // EPI_GELU_BACK: pre-activation in
};

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
  // d/dx x Phi(x) = Phi(x) + x phi(x)
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

template <int MODE>
__global__ void __launch_bounds__(256) mlp_gemm_kernel(GemmArgs g) {
  __shared__ float As[kTileK][kTile + 4];
  __shared__ float Bs[kTileK][kTile + 4];
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int tid = threadIdx.x;
  const int tm = tid >> 4, tn = tid & 15;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;

  for (int k0 = 0; k0 < g.K; k0 += kTileK) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      // A tile 64 (m) x 16 (k)
      const int mm = idx >> 4, kk = idx & 15;
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < g.M && k < g.K) ? g.A[(long)m * g.sam + (long)k * g.sak] : 0.0f;
      // B tile 16 (k) x 64 (n)
      const int kb = idx >> 6, nb = idx & 63;
      const int k2 = k0 + kb, n = n0 + nb;
      Bs[kb][nb] = (k2 < g.K && n < g.N) ? g.B[(long)k2 * g.sbk + (long)n * g.sbn] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTileK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tm * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tn * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + tm * 4 + i;
    if (m >= g.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tn * 4 + j;
      if (n >= g.N) continue;
      const long o = (long)m * g.N + n;
      float v = acc[i][j];
      if (MODE == EPI_BIAS) {
        v += g.bias[n];
      } else if (MODE == EPI_BIAS_GELU) {
        v += g.bias[n];
        g.Z[o] = v;
        v = gelu(v);
      } else if (MODE == EPI_GELU_BACK) {
        v *= gelu_grad(g.Zin[o]);
      }
      g.C[o] = v;
    }
  }
}

// db[n] = sum_m dZ[m][n], summed in row order
__global__ void mlp_colsum_kernel(const float* dZ, int M, int N, float* db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += dZ[(long)m * N + n];
  db[n] = s;
}

// loss = sum_ij w_ij (y_ij - t_ij)^2 / rows ; dY = 2 w (y - t) / rows
__global__ void __launch_bounds__(256) mlp_loss_kernel(const float* Y, const float* T, const float* W, long count,
                                                       int rows, float* dY, double* loss_out) {
  __shared__ double red[256];
  double s = 0.0;
  const float inv = 1.0f / float(rows);
  for (long i = threadIdx.x; i < count; i += 256) {
    const float d = Y[i] - T[i];
    const float w = W ? W[i] : 1.0f;
    s += double(w) * double(d) * double(d);
    if (dY) dY[i] = 2.0f * w * d * inv;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss_out) *loss_out = red[0] / double(rows);
}

__global__ void mlp_adam_kernel(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
                                float eps, float bc1, float bc2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * bc1) / (sqrtf(vi * bc2) + eps);
  }
}

__global__ void mlp_gather_kernel(const float* src, int cols, const int* idx, int rows, float* dst) {
  const long total = (long)rows * cols;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long r = e / cols, c = e - r * cols;
    dst[e] = src[(long)idx[r] * cols + c];
  }
}

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(b, 256));
    if (e == hipSuccess) bytes = b;
    return e;
  }
  float* f() const { return static_cast<float*>(p); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct MlpCtx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_layers = 0;            // linear layers
  int width[kMaxLayers + 1];   // width[0] = inputs, width[n_layers] = outputs
  size_t w_off[kMaxLayers], b_off[kMaxLayers];
  size_t n_params = 0;
  uint64_t adam_t = 0;
  Buf params, grads, adam_m, adam_v;
  Buf act[kMaxLayers + 1];     // act[0] = input rows, act[l] = GELU(Z_l) or outputs
  Buf pre[kMaxLayers];         // Z_l of hidden layers
  Buf dA, dB;                  // backward ping-pong, rows x max width
  Buf data_x, data_t, data_w, bt, bw, idx, loss;
  int cap_rows = 0;
  std::string err;
};

thread_local std::string g_mlp_err;

int mlp_fail(MlpCtx* c, int code, const std::string& m) {
  if (c) c->err = m;
  g_mlp_err = m;
  return code;
}

#define MLPCHK(c, expr)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return mlp_fail(c, TWMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

hipError_t gemm(MlpCtx* c, int mode, const GemmArgs& a) {
  dim3 grid((unsigned)((a.N + kTile - 1) / kTile), (unsigned)((a.M + kTile - 1) / kTile));
  switch (mode) {
    case EPI_STORE: hipLaunchKernelGGL(mlp_gemm_kernel<EPI_STORE>, grid, dim3(256), 0, c->stream, a); break;
    case EPI_BIAS: hipLaunchKernelGGL(mlp_gemm_kernel<EPI_BIAS>, grid, dim3(256), 0, c->stream, a); break;
    case EPI_BIAS_GELU: hipLaunchKernelGGL(mlp_gemm_kernel<EPI_BIAS_GELU>, grid, dim3(256), 0, c->stream, a); break;
    default: hipLaunchKernelGGL(mlp_gemm_kernel<EPI_GELU_BACK>, grid, dim3(256), 0, c->stream, a); break;
  }
  return hipGetLastError();
}

int max_width(const MlpCtx* c) {
  int w = 0;
  for (int l = 0; l <= c->n_layers; ++l) w = std::max(w, c->width[l]);
  return w;
}

int ensure_rows(MlpCtx* c, int rows) {
  if (rows <= c->cap_rows) return TWMC_OK;
  for (int l = 0; l <= c->n_layers; ++l) MLPCHK(c, c->act[l].ensure((size_t)rows * c->width[l] * 4));
  for (int l = 1; l < c->n_layers; ++l) MLPCHK(c, c->pre[l].ensure((size_t)rows * c->width[l] * 4));
  const size_t mw = (size_t)max_width(c);
  MLPCHK(c, c->dA.ensure((size_t)rows * mw * 4));
  MLPCHK(c, c->dB.ensure((size_t)rows * mw * 4));
  MLPCHK(c, c->bt.ensure((size_t)rows * c->width[c->n_layers] * 4));
  MLPCHK(c, c->bw.ensure((size_t)rows * c->width[c->n_layers] * 4));
  c->cap_rows = rows;
  return TWMC_OK;
}

// forward over act[0] (rows x inputs); outputs in act[n_layers]
int forward(MlpCtx* c, int rows) {
  for (int l = 1; l <= c->n_layers; ++l) {
    GemmArgs a{};
    a.M = rows;
    a.N = c->width[l];
    a.K = c->width[l - 1];
    a.A = c->act[l - 1].f();
    a.sam = a.K;
    a.sak = 1;
    a.B = c->params.f() + c->w_off[l - 1];
    a.sbk = a.N;
    a.sbn = 1;
    a.C = c->act[l].f();
    a.bias = c->params.f() + c->b_off[l - 1];
    const bool hidden = l < c->n_layers;
    if (hidden) a.Z = c->pre[l].f();
    MLPCHK(c, gemm(c, hidden ? EPI_BIAS_GELU : EPI_BIAS, a));
  }
  return TWMC_OK;
}

// backward from dY (in dA) to the parameter gradients
int backward(MlpCtx* c, int rows) {
  float* dcur = c->dA.f();
  float* dnext = c->dB.f();
  for (int l = c->n_layers; l >= 1; --l) {
    const int in = c->width[l - 1], out = c->width[l];
    // dW_l = A_{l-1}^T dZ_l : (in x rows) (rows x out)
    GemmArgs w{};
    w.M = in;
    w.N = out;
    w.K = rows;
    w.A = c->act[l - 1].f();
    w.sam = 1;
    w.sak = in;
    w.B = dcur;
    w.sbk = out;
    w.sbn = 1;
    w.C = c->grads.f() + c->w_off[l - 1];
    MLPCHK(c, gemm(c, EPI_STORE, w));
    hipLaunchKernelGGL(mlp_colsum_kernel, dim3((unsigned)((out + 255) / 256)), dim3(256), 0, c->stream,
                       (const float*)dcur, rows, out, c->grads.f() + c->b_off[l - 1]);
    MLPCHK(c, hipGetLastError());
    if (l > 1) {
      // dZ_{l-1} = (dZ_l W_l^T) . GELU'(Z_{l-1}) : (rows x out) (out x in)
      GemmArgs d{};
      d.M = rows;
      d.N = in;
      d.K = out;
      d.A = dcur;
      d.sam = out;
      d.sak = 1;
      d.B = c->params.f() + c->w_off[l - 1];
      d.sbk = 1;
      d.sbn = out;
      d.C = dnext;
      d.Zin = c->pre[l - 1].f();
      MLPCHK(c, gemm(c, EPI_GELU_BACK, d));
      std::swap(dcur, dnext);
    }
  }
  return TWMC_OK;
}

int loss_and_grad(MlpCtx* c, const float* T, const float* W, int rows, double* loss_dev) {
  const int out = c->width[c->n_layers];
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, c->stream, (const float*)c->act[c->n_layers].f(), T, W,
                     (long)rows * out, rows, c->dA.f(), loss_dev);
  MLPCHK(c, hipGetLastError());
  return backward(c, rows);
}

// splitmix64: host-side shuffles and initial weights
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace
}  // namespace twmc

using namespace twmc;

extern "C" {

const char* twmc_mlp_last_error(const twmc_mlp_ctx* ctx) {
  if (ctx) return reinterpret_cast<const MlpCtx*>(ctx)->err.c_str();
  return g_mlp_err.c_str();
}

int twmc_mlp_create(const twmc_mlp_config* cfg, int device, twmc_mlp_ctx** ctx_out) {
  if (!cfg || !ctx_out) return mlp_fail(nullptr, TWMC_ERR_INVALID, "null argument");
  *ctx_out = nullptr;
  if (cfg->n_in < 1 || cfg->n_out < 1 || cfg->n_hidden < 1 || cfg->n_hidden_layers < 0 ||
      cfg->n_hidden_layers + 1 > kMaxLayers || cfg->n_in > 65536 || cfg->n_out > 65536 || cfg->n_hidden > 65536)
    return mlp_fail(nullptr, TWMC_ERR_INVALID, "bad network shape");
  MlpCtx* c = new MlpCtx();
  c->device = device;
  c->n_layers = cfg->n_hidden_layers + 1;
  c->width[0] = cfg->n_in;
  for (int l = 1; l < c->n_layers; ++l) c->width[l] = cfg->n_hidden;
  c->width[c->n_layers] = cfg->n_out;
  size_t off = 0;
  for (int l = 0; l < c->n_layers; ++l) {
    c->w_off[l] = off;
    off += (size_t)c->width[l] * c->width[l + 1];
    c->b_off[l] = off;
    off += (size_t)c->width[l + 1];
  }
  c->n_params = off;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = c->params.ensure(off * 4);
  if (e == hipSuccess) e = c->grads.ensure(off * 4);
  if (e == hipSuccess) e = c->adam_m.ensure(off * 4);
  if (e == hipSuccess) e = c->adam_v.ensure(off * 4);
  if (e != hipSuccess) {
    std::string m = std::string("twmc_mlp_create: ") + hipGetErrorString(e);
    twmc_mlp_destroy(reinterpret_cast<twmc_mlp_ctx*>(c));
    return mlp_fail(nullptr, TWMC_ERR_HIP, m);
  }
  // Glorot-uniform weights, zero biases, deterministic in cfg->seed
  std::vector<float> p(off, 0.0f);
  uint64_t s = cfg->seed;
  for (int l = 0; l < c->n_layers; ++l) {
    const double a = std::sqrt(6.0 / double(c->width[l] + c->width[l + 1]));
    const size_t n = (size_t)c->width[l] * c->width[l + 1];
    for (size_t i = 0; i < n; ++i) {
      const double u = double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      p[c->w_off[l] + i] = float((2.0 * u - 1.0) * a);
    }
  }
  *ctx_out = reinterpret_cast<twmc_mlp_ctx*>(c);
  int rc = twmc_mlp_set_params(*ctx_out, p.data());
  if (rc != TWMC_OK) {
    std::string m = c->err;
    twmc_mlp_destroy(reinterpret_cast<twmc_mlp_ctx*>(c));
    *ctx_out = nullptr;
    return mlp_fail(nullptr, rc, m);
  }
  return TWMC_OK;
}

void twmc_mlp_destroy(twmc_mlp_ctx* ctx) {
  if (!ctx) return;
  MlpCtx* c = reinterpret_cast<MlpCtx*>(ctx);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  Buf* bufs[] = {&c->params, &c->grads, &c->adam_m, &c->adam_v, &c->dA, &c->dB, &c->data_x, &c->data_t,
                 &c->data_w, &c->bt, &c->bw, &c->idx, &c->loss};
  for (Buf* b : bufs) b->release();
  for (int l = 0; l <= kMaxLayers; ++l) c->act[l].release();
  for (int l = 0; l < kMaxLayers; ++l) c->pre[l].release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

uint64_t twmc_mlp_n_params(const twmc_mlp_ctx* ctx) { return ctx ? reinterpret_cast<const MlpCtx*>(ctx)->n_params : 0; }

int twmc_mlp_get_params(twmc_mlp_ctx* ctx, float* out) {
  if (!ctx || !out) return mlp_fail(nullptr, TWMC_ERR_INVALID, "null argument");
  MlpCtx* c = reinterpret_cast<MlpCtx*>(ctx);
  MLPCHK(c, hipSetDevice(c->device));
  MLPCHK(c, hipStreamSynchronize(c->stream));
  MLPCHK(c, hipMemcpy(out, c->params.p, c->n_params * 4, hipMemcpyDeviceToHost));
  return TWMC_OK;
}

int twmc_mlp_set_params(twmc_mlp_ctx* ctx, const float* in) {
  if (!ctx || !in) return mlp_fail(nullptr, TWMC_ERR_INVALID, "null argument");
  MlpCtx* c = reinterpret_cast<MlpCtx*>(ctx);
  MLPCHK(c, hipSetDevice(c->device));
  MLPCHK(c, hipStreamSynchronize(c->stream));
  MLPCHK(c, hipMemcpy(c->params.p, in, c->n_params * 4, hipMemcpyHostToDevice));
  std::vector<float> zero(c->n_params, 0.0f);
  MLPCHK(c, hipMemcpy(c->adam_m.p, zero.data(), c->n_params * 4, hipMemcpyHostToDevice));
  MLPCHK(c, hipMemcpy(c->adam_v.p, zero.data(), c->n_params * 4, hipMemcpyHostToDevice));
  c->adam_t = 0;
  return TWMC_OK;
}

int twmc_mlp_forward(twmc_mlp_ctx* ctx, const float* x, uint64_t rows, float* y) {
  if (!ctx || !x || !y) return mlp_fail(nullptr, TWMC_ERR_INVALID, "null argument");
  MlpCtx* c = reinterpret_cast<MlpCtx*>(ctx);
  if (rows == 0 || rows > (1u << 24)) return mlp_fail(c, TWMC_ERR_INVALID, "rows out of range");
  MLPCHK(c, hipSetDevice(c->device));
  int rc = ensure_rows(c, (int)rows);
  if (rc != TWMC_OK) return rc;
  MLPCHK(c, hipMemcpyAsync(c->act[0].p, x, rows * c->width[0] * 4, hipMemcpyHostToDevice, c->stream));
  rc = forward(c, (int)rows);
  if (rc != TWMC_OK) return rc;
  MLPCHK(c, hipMemcpyAsync(y, c->act[c->n_layers].p, rows * c->width[c->n_layers] * 4, hipMemcpyDeviceToHost,
                           c->stream));
  MLPCHK(c, hipStreamSynchronize(c->stream));
  return TWMC_OK;
}

int twmc_mlp_loss_grad(twmc_mlp_ctx* ctx, const float* x, const float* t, const float* w, uint64_t rows, double* loss,
                       float* grad) {
  if (!ctx || !x || !t) return mlp_fail(nullptr, TWMC_ERR_INVALID, "null argument");
  MlpCtx* c = reinterpret_cast<MlpCtx*>(ctx);
  if (rows == 0 || rows > (1u << 24)) return mlp_fail(c, TWMC_ERR_INVALID, "rows out of range");
  MLPCHK(c, hipSetDevice(c->device));
  int rc = ensure_rows(c, (int)rows);
  if (rc != TWMC_OK) return rc;
  const size_t out = (size_t)c->width[c->n_layers];
  MLPCHK(c, c->loss.ensure(sizeof(double)));
  MLPCHK(c, hipMemcpyAsync(c->act[0].p, x, rows * c->width[0] * 4, hipMemcpyHostToDevice, c->stream));
  MLPCHK(c, hipMemcpyAsync(c->bt.p, t, rows * out * 4, hipMemcpyHostToDevice, c->stream));
  if (w) MLPCHK(c, hipMemcpyAsync(c->bw.p, w, rows * out * 4, hipMemcpyHostToDevice, c->stream));
  rc = forward(c, (int)rows);
  if (rc == TWMC_OK)
    rc = loss_and_grad(c, c->bt.f(), w ? c->bw.f() : nullptr, (int)rows, static_cast<double*>(c->loss.p));
  if (rc != TWMC_OK) return rc;
  double l = 0.0;
  MLPCHK(c, hipMemcpyAsync(&l, c->loss.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (grad) MLPCHK(c, hipMemcpyAsync(grad, c->grads.p, c->n_params * 4, hipMemcpyDeviceToHost, c->stream));
  MLPCHK(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = l;
  return TWMC_OK;
}

int twmc_mlp_train(twmc_mlp_ctx* ctx, const float* x, const float* t, const float* w, uint64_t rows,
                   const twmc_mlp_train_opts* o, double* epoch_loss) {
  if (!ctx || !x || !t || !o) return mlp_fail(nullptr, TWMC_ERR_INVALID, "null argument");
  MlpCtx* c = reinterpret_cast<MlpCtx*>(ctx);
  if (rows == 0 || rows > (1u << 24)) return mlp_fail(c, TWMC_ERR_INVALID, "rows out of range");
  if (o->epochs < 0 || o->batch < 1 || !(o->lr > 0.0f) || !(o->beta1 >= 0.0f && o->beta1 < 1.0f) ||
      !(o->beta2 >= 0.0f && o->beta2 < 1.0f) || !(o->eps > 0.0f))
    return mlp_fail(c, TWMC_ERR_INVALID, "bad training options");
  MLPCHK(c, hipSetDevice(c->device));
  const int nin = c->width[0], nout = c->width[c->n_layers];
  const int batch = (int)std::min<uint64_t>((uint64_t)o->batch, rows);
  int rc = ensure_rows(c, batch);
  if (rc != TWMC_OK) return rc;
  MLPCHK(c, c->data_x.ensure(rows * nin * 4));
  MLPCHK(c, c->data_t.ensure(rows * nout * 4));
  if (w) MLPCHK(c, c->data_w.ensure(rows * nout * 4));
  MLPCHK(c, c->idx.ensure(rows * 4));
  const uint64_t n_batches = rows / (uint64_t)batch;  // the remainder of a shuffled epoch is dropped
  MLPCHK(c, c->loss.ensure(std::max<uint64_t>(n_batches, 1) * sizeof(double)));
  MLPCHK(c, hipMemcpyAsync(c->data_x.p, x, rows * nin * 4, hipMemcpyHostToDevice, c->stream));
  MLPCHK(c, hipMemcpyAsync(c->data_t.p, t, rows * nout * 4, hipMemcpyHostToDevice, c->stream));
  if (w) MLPCHK(c, hipMemcpyAsync(c->data_w.p, w, rows * nout * 4, hipMemcpyHostToDevice, c->stream));
  std::vector<int> perm(rows);
  std::vector<double> bl(n_batches);
  uint64_t s = o->shuffle_seed;
  for (int ep = 0; ep < o->epochs; ++ep) {
    for (uint64_t i = 0; i < rows; ++i) perm[i] = (int)i;
    for (uint64_t i = rows - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(s) % (i + 1)]);
    // the copy must finish before perm is refilled next epoch
    MLPCHK(c, hipMemcpyAsync(c->idx.p, perm.data(), rows * 4, hipMemcpyHostToDevice, c->stream));
    for (uint64_t b = 0; b < n_batches; ++b) {
      const int* ib = static_cast<const int*>(c->idx.p) + b * batch;
      const int gx = (int)std::min<int64_t>(((int64_t)batch * nout + 255) / 256, 4096);
      hipLaunchKernelGGL(mlp_gather_kernel, dim3((unsigned)gx), dim3(256), 0, c->stream, (const float*)c->data_x.f(),
                         nin, ib, batch, c->act[0].f());
      hipLaunchKernelGGL(mlp_gather_kernel, dim3((unsigned)gx), dim3(256), 0, c->stream, (const float*)c->data_t.f(),
                         nout, ib, batch, c->bt.f());
      if (w)
        hipLaunchKernelGGL(mlp_gather_kernel, dim3((unsigned)gx), dim3(256), 0, c->stream,
                           (const float*)c->data_w.f(), nout, ib, batch, c->bw.f());
      MLPCHK(c, hipGetLastError());
      rc = forward(c, batch);
      if (rc == TWMC_OK)
        rc = loss_and_grad(c, c->bt.f(), w ? c->bw.f() : nullptr, batch, static_cast<double*>(c->loss.p) + b);
      if (rc != TWMC_OK) return rc;
      c->adam_t += 1;
      const double bc1 = 1.0 / (1.0 - std::pow((double)o->beta1, (double)c->adam_t));
      const double bc2 = 1.0 / (1.0 - std::pow((double)o->beta2, (double)c->adam_t));
      const int ga = (int)std::min<size_t>((c->n_params + 255) / 256, 2048);
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)ga), dim3(256), 0, c->stream, c->params.f(),
                         (const float*)c->grads.f(), c->adam_m.f(), c->adam_v.f(), (long)c->n_params, o->lr,
                         o->beta1, o->beta2, o->eps, (float)bc1, (float)bc2);
      MLPCHK(c, hipGetLastError());
    }
    if (n_batches) MLPCHK(c, hipMemcpyAsync(bl.data(), c->loss.p, n_batches * 8, hipMemcpyDeviceToHost, c->stream));
    MLPCHK(c, hipStreamSynchronize(c->stream));
    if (epoch_loss) {
      double m = 0.0;
      for (double v : bl) m += v;
      epoch_loss[ep] = n_batches ? m / double(n_batches) : 0.0;
    }
  }
  return TWMC_OK;
}

}  // extern "C"
