// mlp.hip - reflectance surrogate: a multilayer perceptron trained and served on gfx950 with
// hand-written BF16 matrix-core kernels (v_mfma_f32_32x32x8_bf16, FP32 accumulation), GELU hidden
// layers, mean-squared-error loss and Adam with FP32 master weights.
//
// Layout: weights W_l are [in_l x out_l] row-major, activations [batch x width] row-major, all
// matrix operands BF16 (raw uint16 bits), accumulators and pre-activations FP32.
//
// One GEMM kernel serves every product of the forward and backward passes:
//   forward   Z = X W + b          A = X (batch x in),   B = W
//   backward  dA_prev = dZ W^T     A = dZ,               B = W^T (transposed operand)
//             dW = X^T dZ          A = X^T (transposed), B = dZ
// Tiles of 64x64 outputs per 256-thread workgroup (4 waves, one 32x32 MFMA tile each), K staged
// through LDS in slices of 32 with zero fill at the edges.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/ymc3d.h"
#include "host_util.h"

namespace ymc {
namespace mlp {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * __expf(-0.5f * z * z);
  return cdf + z * pdf;
}

enum Epilogue { kStoreF32 = 0, kBiasGelu = 1, kBias = 2 };

constexpr int kTile = 64, kSlice = 32, kThreads = 256;

// C[M x N] = op(A)[M x K] * op(B)[K x N].
// transA: op(A)(m, k) = A[k * lda + m], else A[m * lda + k].
// transB: op(B)(k, n) = B[n * ldb + k], else B[k * ldb + n].
template <int EPI>
__global__ __launch_bounds__(kThreads) void gemm_bf16(const uint16_t *__restrict__ A, int lda, int transA,
                                                      const uint16_t *__restrict__ B, int ldb, int transB, int M,
                                                      int N, int K, const float *__restrict__ bias,
                                                      float *__restrict__ outF, int ldo, uint16_t *__restrict__ outH,
                                                      int ldh) {
  __shared__ uint16_t sA[kTile][kSlice + 4];  // [m][k]
  __shared__ uint16_t sB[kTile][kSlice + 4];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  float16_t acc;
  for (int v = 0; v < 16; ++v) acc[v] = 0.0f;

  for (int k0 = 0; k0 < K; k0 += kSlice) {
    // stage 64x32 of op(A) and 32x64 of op(B): 2048 elements each, 8 per thread
    for (int e = tid; e < kTile * kSlice; e += kThreads) {
      int m, k;
      if (transA) {  // consecutive threads walk m (contiguous in memory)
        m = e % kTile;
        k = e / kTile;
      } else {
        k = e % kSlice;
        m = e / kSlice;
      }
      const int gm = m0 + m;
      int gk = k0 + k;
      uint16_t v = 0;
      if (gm < M && gk < K) v = transA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      sA[m][k] = v;
      int n;
      if (transB) {  // consecutive threads walk k
        k = e % kSlice;
        n = e / kSlice;
      } else {
        n = e % kTile;
        k = e / kTile;
      }
      const int gn = n0 + n;
      gk = k0 + k;
      uint16_t w = 0;
      if (gn < N && gk < K) w = transB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      sB[n][k] = w;
    }
    __syncthreads();
    // four 32x32x8 MFMAs per slice.  Operand lanes: lane l holds row/column l % 32 and the four
    // k values 4*(l/32) .. 4*(l/32)+3 of the 8-deep step.
    const int r = lane & 31, kq = (lane >> 5) * 4;
    for (int kk = 0; kk < kSlice; kk += 8) {
      short4_t a, b;
      for (int q = 0; q < 4; ++q) {
        a[q] = (short)sA[wm + r][kk + kq + q];
        b[q] = (short)sB[wn + r][kk + kq + q];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // Accumulator lanes: value v of lane l is row 8*(v/4) + 4*(l/32) + v%4, column l % 32.
  const int col = n0 + wn + (lane & 31);
  for (int v = 0; v < 16; ++v) {
    const int row = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    if (row >= M || col >= N) continue;
    float z = acc[v];
    if (EPI == kStoreF32) {
      outF[(size_t)row * ldo + col] = z;
    } else {
      z += bias[col];
      outF[(size_t)row * ldo + col] = z;
      if (EPI == kBiasGelu) outH[(size_t)row * ldh + col] = f2bf(gelu(z));
    }
  }
}

// dZ = dA * gelu'(Z) (FP32), rounded to BF16 for the next products.
__global__ void gelu_backward(const float *dA, const float *Z, float *dZ, uint16_t *dZh, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = dA[i] * gelu_grad(Z[i]);
  dZ[i] = g;
  dZh[i] = f2bf(g);
}

// Mean-squared-error gradient: dOut = 2 (out - y) / (rows * cols).  Loss partial sums per block.
__global__ void mse_grad(const float *out, const float *y, float *dOut, uint16_t *dOuth, int n, float scale,
                         float *loss_parts) {
  __shared__ float red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float d = out[i] - y[i];
    dOut[i] = 2.0f * d * scale;
    dOuth[i] = f2bf(2.0f * d * scale);
    l = d * d * scale;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_parts[blockIdx.x] = red[0];
}

// db[c] = sum over rows of dZ[r][c] (FP32).
__global__ void column_sum(const float *dZ, int rows, int cols, float *db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int r = 0; r < rows; ++r) s += dZ[(size_t)r * cols + c];
  db[c] = s;
}

// Adam on FP32 master parameters, refreshing the BF16 weight copy.
__global__ void adam_update(float *p, const float *g, float *m, float *v, uint16_t *ph, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i] = pi;
  if (ph) ph[i] = f2bf(pi);
}

__global__ void to_bf16(const float *x, uint16_t *h, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = f2bf(x[i]);
}

// Gathers rows idx[r] of X (FP32, cols_x) and Y (FP32, cols_y) into a minibatch.
__global__ void gather_rows(const float *X, const float *Y, const int *idx, int rows, int cx, int cy, uint16_t *xb,
                            float *yb) {
  const int r = blockIdx.x, t = threadIdx.x;
  if (r >= rows) return;
  const size_t s = (size_t)idx[r];
  for (int c = t; c < cx; c += blockDim.x) xb[(size_t)r * cx + c] = f2bf(X[s * cx + c]);
  for (int c = t; c < cy; c += blockDim.x) yb[(size_t)r * cy + c] = Y[s * cy + c];
}

}  // namespace mlp
}  // namespace ymc

// ---------------------------------------------------------------------------------------------
// Host side.

using namespace ymc;
using namespace ymc::mlp;

#define MLP_TRY(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t err_ = (expr);                                                                           \
    if (err_ != hipSuccess) return fail(YMC_ERR_DEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(err_), \
                                        __FILE__, __LINE__);                                            \
  } while (0)

struct ymc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<int> widths;           // layer widths, input first
  std::vector<size_t> w_off, b_off;  // offsets of W_l and b_l in the flat parameter vector
  size_t n_params = 0;
  int cap = 0;                       // minibatch capacity
  // device buffers
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 master params, grads, moments
  uint16_t *ph = nullptr;                                         // BF16 copy of the parameters
  std::vector<float *> Z;             // pre-activations per layer [cap x out]
  std::vector<uint16_t *> H;          // BF16 activations per layer (H[0] = input) [cap x width]
  float *dA = nullptr, *dZ = nullptr, *yb = nullptr, *loss_parts = nullptr;
  uint16_t *dZh = nullptr;
  float *X = nullptr, *Y = nullptr;   // training set
  int64_t n_data = 0;
  int *idx = nullptr;
  int64_t adam_t = 0;
};

namespace {

template <class T>
ymc_status dalloc(T **p, size_t count) {
  if (hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return fail(YMC_ERR_NOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
  }
  return YMC_OK;
}

int max_width(const ymc_mlp *n) { return *std::max_element(n->widths.begin(), n->widths.end()); }

ymc_status ensure_capacity(ymc_mlp *n, int batch) {
  if (batch <= n->cap) return YMC_OK;
  const int L = (int)n->widths.size() - 1;
  for (float *z : n->Z) (void)hipFree(z);
  for (uint16_t *h : n->H) (void)hipFree(h);
  (void)hipFree(n->dA);
  (void)hipFree(n->dZ);
  (void)hipFree(n->dZh);
  (void)hipFree(n->yb);
  (void)hipFree(n->loss_parts);
  n->Z.assign(L, nullptr);
  n->H.assign(L + 1, nullptr);
  ymc_status st;
  const size_t wmax = (size_t)max_width(n);
  for (int l = 0; l < L; ++l)
    if ((st = dalloc(&n->Z[l], (size_t)batch * n->widths[l + 1]))) return st;
  for (int l = 0; l <= L; ++l)
    if ((st = dalloc(&n->H[l], (size_t)batch * n->widths[l]))) return st;
  if ((st = dalloc(&n->dA, batch * wmax)) || (st = dalloc(&n->dZ, batch * wmax)) ||
      (st = dalloc(&n->dZh, batch * wmax)) || (st = dalloc(&n->yb, (size_t)batch * n->widths[L])) ||
      (st = dalloc(&n->loss_parts, (size_t)batch * n->widths[L] / 256 + 2)))
    return st;
  n->cap = batch;
  return YMC_OK;
}

void gemm(hipStream_t s, int epi, const uint16_t *A, int lda, int tA, const uint16_t *B, int ldb, int tB, int M, int N,
          int K, const float *bias, float *outF, int ldo, uint16_t *outH, int ldh) {
  dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  if (epi == kStoreF32)
    hipLaunchKernelGGL(gemm_bf16<kStoreF32>, grid, dim3(kThreads), 0, s, A, lda, tA, B, ldb, tB, M, N, K, bias, outF,
                       ldo, outH, ldh);
  else if (epi == kBiasGelu)
    hipLaunchKernelGGL(gemm_bf16<kBiasGelu>, grid, dim3(kThreads), 0, s, A, lda, tA, B, ldb, tB, M, N, K, bias, outF,
                       ldo, outH, ldh);
  else
    hipLaunchKernelGGL(gemm_bf16<kBias>, grid, dim3(kThreads), 0, s, A, lda, tA, B, ldb, tB, M, N, K, bias, outF, ldo,
                       outH, ldh);
}

// Forward pass of `batch` rows already in H[0]; output pre-activations in Z[L-1].
void forward(ymc_mlp *n, int batch) {
  const int L = (int)n->widths.size() - 1;
  for (int l = 0; l < L; ++l) {
    const int in = n->widths[l], out = n->widths[l + 1];
    gemm(n->stream, l + 1 < L ? kBiasGelu : kBias, n->H[l], in, 0, n->ph + n->w_off[l], out, 0, batch, out, in,
         n->p + n->b_off[l], n->Z[l], out, n->H[l + 1], out);
  }
}

// Backward pass for targets in yb; gradients into g; returns the loss via loss_parts.
int backward(ymc_mlp *n, int batch) {
  const int L = (int)n->widths.size() - 1;
  const int out = n->widths[L];
  const int cnt = batch * out;
  const int blocks = (cnt + 255) / 256;
  hipLaunchKernelGGL(mse_grad, dim3(blocks), dim3(256), 0, n->stream, n->Z[L - 1], n->yb, n->dZ, n->dZh, cnt,
                     1.0f / (float)cnt, n->loss_parts);
  for (int l = L - 1; l >= 0; --l) {
    const int in = n->widths[l], ow = n->widths[l + 1];
    // bias gradient from the FP32 dZ, weight gradient dW = H[l]^T dZ
    hipLaunchKernelGGL(column_sum, dim3((ow + 255) / 256), dim3(256), 0, n->stream, n->dZ, batch, ow,
                       n->g + n->b_off[l]);
    gemm(n->stream, kStoreF32, n->H[l], in, 1, n->dZh, ow, 0, in, ow, batch, nullptr, n->g + n->w_off[l], ow, nullptr,
         0);
    if (l == 0) break;
    // dA_prev = dZ W^T, then dZ_prev = dA_prev * gelu'(Z_prev)
    gemm(n->stream, kStoreF32, n->dZh, ow, 0, n->ph + n->w_off[l], ow, 1, batch, in, ow, nullptr, n->dA, in, nullptr, 0);
    const int c2 = batch * in;
    hipLaunchKernelGGL(gelu_backward, dim3((c2 + 255) / 256), dim3(256), 0, n->stream, n->dA, n->Z[l - 1], n->dZ,
                       n->dZh, c2);
  }
  return blocks;
}

ymc_status read_loss(ymc_mlp *n, int parts, float *loss) {
  std::vector<float> h(parts);
  MLP_TRY(hipMemcpyAsync(h.data(), n->loss_parts, parts * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  MLP_TRY(hipStreamSynchronize(n->stream));
  double s = 0;
  for (float x : h) s += x;
  *loss = (float)s;
  return YMC_OK;
}

}  // namespace

extern "C" {

ymc_status ymc_mlp_create(int32_t device, const int32_t *widths, int32_t n_widths, uint64_t seed, ymc_mlp **out) {
  if (!out || !widths || n_widths < 2) return fail(YMC_ERR_ARG, "ymc_mlp_create: need >= 2 widths");
  for (int i = 0; i < n_widths; ++i)
    if (widths[i] < 1 || widths[i] > 8192) return fail(YMC_ERR_ARG, "ymc_mlp_create: width %d out of range", widths[i]);
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return fail(YMC_ERR_DEVICE, "hipSetDevice(%d) failed", device);
  ymc_mlp *n = new ymc_mlp();
  n->device = device;
  n->widths.assign(widths, widths + n_widths);
  const int L = n_widths - 1;
  for (int l = 0; l < L; ++l) {
    n->w_off.push_back(n->n_params);
    n->n_params += (size_t)widths[l] * widths[l + 1];
    n->b_off.push_back(n->n_params);
    n->n_params += widths[l + 1];
  }
  ymc_status st;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
      (st = dalloc(&n->p, n->n_params)) || (st = dalloc(&n->g, n->n_params)) || (st = dalloc(&n->m, n->n_params)) ||
      (st = dalloc(&n->v, n->n_params)) || (st = dalloc(&n->ph, n->n_params))) {
    ymc_mlp_destroy(n);
    return fail(YMC_ERR_NOMEM, "ymc_mlp_create: allocation failed");
  }
  // He-style initialisation from a seeded host generator; biases zero
  std::vector<float> init(n->n_params, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < L; ++l) {
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / widths[l]));
    for (size_t q = 0; q < (size_t)widths[l] * widths[l + 1]; ++q) init[n->w_off[l] + q] = nd(rng);
  }
  if ((st = ymc_mlp_set_params(n, init.data(), n->n_params))) {
    ymc_mlp_destroy(n);
    return st;
  }
  *out = n;
  return YMC_OK;
}

void ymc_mlp_destroy(ymc_mlp *n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void *q : {(void *)n->p, (void *)n->g, (void *)n->m, (void *)n->v, (void *)n->ph, (void *)n->dA, (void *)n->dZ,
                  (void *)n->dZh, (void *)n->yb, (void *)n->loss_parts, (void *)n->X, (void *)n->Y, (void *)n->idx})
    if (q) (void)hipFree(q);
  for (float *z : n->Z) (void)hipFree(z);
  for (uint16_t *h : n->H) (void)hipFree(h);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

uint64_t ymc_mlp_num_params(const ymc_mlp *n) { return n ? n->n_params : 0; }

ymc_status ymc_mlp_set_params(ymc_mlp *n, const float *params, uint64_t count) {
  if (!n || !params || count != n->n_params) return fail(YMC_ERR_ARG, "ymc_mlp_set_params: expected %zu values", n ? n->n_params : 0);
  MLP_TRY(hipSetDevice(n->device));
  MLP_TRY(hipMemcpy(n->p, params, count * sizeof(float), hipMemcpyHostToDevice));
  MLP_TRY(hipMemset(n->m, 0, count * sizeof(float)));
  MLP_TRY(hipMemset(n->v, 0, count * sizeof(float)));
  hipLaunchKernelGGL(to_bf16, dim3((count + 255) / 256), dim3(256), 0, n->stream, n->p, n->ph, (int)count);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipStreamSynchronize(n->stream));
  n->adam_t = 0;
  return YMC_OK;
}

ymc_status ymc_mlp_get_params(ymc_mlp *n, float *params, uint64_t count) {
  if (!n || !params || count != n->n_params) return fail(YMC_ERR_ARG, "ymc_mlp_get_params: expected %zu values", n ? n->n_params : 0);
  MLP_TRY(hipSetDevice(n->device));
  MLP_TRY(hipStreamSynchronize(n->stream));
  MLP_TRY(hipMemcpy(params, n->p, count * sizeof(float), hipMemcpyDeviceToHost));
  return YMC_OK;
}

ymc_status ymc_mlp_forward(ymc_mlp *n, const float *x, int64_t rows, float *y) {
  if (!n || !x || !y || rows < 0) return fail(YMC_ERR_ARG, "ymc_mlp_forward: bad arguments");
  MLP_TRY(hipSetDevice(n->device));
  const int L = (int)n->widths.size() - 1;
  const int in = n->widths[0], out = n->widths[L];
  const int chunk = 4096;
  ymc_status st;
  if ((st = ensure_capacity(n, chunk))) return st;
  float *xd = nullptr;
  if ((st = dalloc(&xd, (size_t)chunk * in))) return st;
  for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
    const int b = (int)std::min<int64_t>(chunk, rows - r0);
    MLP_TRY(hipMemcpyAsync(xd, x + r0 * in, (size_t)b * in * sizeof(float), hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(to_bf16, dim3((b * in + 255) / 256), dim3(256), 0, n->stream, xd, n->H[0], b * in);
    forward(n, b);
    MLP_TRY(hipMemcpyAsync(y + r0 * out, n->Z[L - 1], (size_t)b * out * sizeof(float), hipMemcpyDeviceToHost,
                           n->stream));
  }
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipStreamSynchronize(n->stream));
  (void)hipFree(xd);
  return YMC_OK;
}

ymc_status ymc_mlp_time_forward(ymc_mlp *n, int32_t rows, int32_t reps, double *ms_per_pass) {
  if (!n || !ms_per_pass || rows < 1 || reps < 1) return fail(YMC_ERR_ARG, "ymc_mlp_time_forward: bad arguments");
  MLP_TRY(hipSetDevice(n->device));
  ymc_status st;
  if ((st = ensure_capacity(n, rows))) return st;
  MLP_TRY(hipMemsetAsync(n->H[0], 0, (size_t)rows * n->widths[0] * sizeof(uint16_t), n->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MLP_TRY(hipEventCreate(&e0));
  MLP_TRY(hipEventCreate(&e1));
  forward(n, rows);  // warm-up
  MLP_TRY(hipEventRecord(e0, n->stream));
  for (int r = 0; r < reps; ++r) forward(n, rows);
  MLP_TRY(hipEventRecord(e1, n->stream));
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  MLP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_pass = (double)ms / reps;
  return YMC_OK;
}

ymc_status ymc_mlp_gradients(ymc_mlp *n, const float *x, const float *y, int32_t rows, float *grad, float *loss) {
  if (!n || !x || !y || !grad || !loss || rows < 1) return fail(YMC_ERR_ARG, "ymc_mlp_gradients: bad arguments");
  MLP_TRY(hipSetDevice(n->device));
  const int L = (int)n->widths.size() - 1;
  const int in = n->widths[0], out = n->widths[L];
  ymc_status st;
  if ((st = ensure_capacity(n, rows))) return st;
  float *xd = nullptr;
  if ((st = dalloc(&xd, (size_t)rows * in))) return st;
  MLP_TRY(hipMemcpyAsync(xd, x, (size_t)rows * in * sizeof(float), hipMemcpyHostToDevice, n->stream));
  MLP_TRY(hipMemcpyAsync(n->yb, y, (size_t)rows * out * sizeof(float), hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(to_bf16, dim3((rows * in + 255) / 256), dim3(256), 0, n->stream, xd, n->H[0], rows * in);
  forward(n, rows);
  const int parts = backward(n, rows);
  MLP_TRY(hipGetLastError());
  if ((st = read_loss(n, parts, loss))) return st;
  MLP_TRY(hipMemcpy(grad, n->g, n->n_params * sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(xd);
  return YMC_OK;
}

ymc_status ymc_mlp_set_data(ymc_mlp *n, const float *x, const float *y, int64_t rows) {
  if (!n || !x || !y || rows < 1) return fail(YMC_ERR_ARG, "ymc_mlp_set_data: bad arguments");
  MLP_TRY(hipSetDevice(n->device));
  const int L = (int)n->widths.size() - 1;
  (void)hipFree(n->X);
  (void)hipFree(n->Y);
  (void)hipFree(n->idx);
  n->X = nullptr;
  n->Y = nullptr;
  n->idx = nullptr;
  ymc_status st;
  if ((st = dalloc(&n->X, (size_t)rows * n->widths[0])) || (st = dalloc(&n->Y, (size_t)rows * n->widths[L])) ||
      (st = dalloc(&n->idx, (size_t)rows)))
    return st;
  MLP_TRY(hipMemcpy(n->X, x, (size_t)rows * n->widths[0] * sizeof(float), hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(n->Y, y, (size_t)rows * n->widths[L] * sizeof(float), hipMemcpyHostToDevice));
  n->n_data = rows;
  return YMC_OK;
}

ymc_status ymc_mlp_train(ymc_mlp *n, int32_t epochs, int32_t batch, float lr, float lr_final, uint64_t seed,
                         float *epoch_loss) {
  if (!n || !n->X || epochs < 1 || batch < 1) return fail(YMC_ERR_ARG, "ymc_mlp_train: set data first; epochs, batch >= 1");
  MLP_TRY(hipSetDevice(n->device));
  ymc_status st;
  if ((st = ensure_capacity(n, batch))) return st;
  const int L = (int)n->widths.size() - 1;
  const int64_t steps_per_epoch = std::max<int64_t>(1, n->n_data / batch);
  const int64_t total = steps_per_epoch * epochs;
  std::mt19937_64 rng(seed);
  std::vector<int> perm(n->n_data);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = 0; i < n->n_data; ++i) perm[i] = (int)i;
    std::shuffle(perm.begin(), perm.end(), rng);
    MLP_TRY(hipMemcpyAsync(n->idx, perm.data(), n->n_data * sizeof(int), hipMemcpyHostToDevice, n->stream));
    double loss_sum = 0;
    for (int64_t s = 0; s < steps_per_epoch; ++s) {
      hipLaunchKernelGGL(gather_rows, dim3(batch), dim3(64), 0, n->stream, n->X, n->Y, n->idx + s * batch, batch,
                         n->widths[0], n->widths[L], n->H[0], n->yb);
      forward(n, batch);
      const int parts = backward(n, batch);
      // cosine-decayed learning rate
      const int64_t t = (int64_t)ep * steps_per_epoch + s;
      const float lr_t = lr_final + 0.5f * (lr - lr_final) * (1.0f + std::cos(3.14159265f * (float)t / (float)total));
      n->adam_t++;
      const float c1 = 1.0f / (1.0f - std::pow(b1, (float)n->adam_t));
      const float c2 = 1.0f / (1.0f - std::pow(b2, (float)n->adam_t));
      hipLaunchKernelGGL(adam_update, dim3((n->n_params + 255) / 256), dim3(256), 0, n->stream, n->p, n->g, n->m, n->v,
                         n->ph, (int)n->n_params, lr_t, b1, b2, eps, c1, c2);
      if (s == steps_per_epoch - 1) {
        float l = 0;
        if ((st = read_loss(n, parts, &l))) return st;
        loss_sum = l;
      }
    }
    MLP_TRY(hipGetLastError());
    if (epoch_loss) epoch_loss[ep] = (float)loss_sum;
  }
  MLP_TRY(hipStreamSynchronize(n->stream));
  return YMC_OK;
}

}  // extern "C"
