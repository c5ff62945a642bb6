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
#include <mutex>
#include <random>
#include <vector>

#include "../../include/ymc3d.h"
#include "host_util.h"

namespace ymc {
namespace mlp {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// FP32 -> BF16, round to nearest even: gfx950's v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// GELU(z) = z Phi(z), Phi the standard normal CDF.  With q(z) = erfc(|z|/sqrt 2) / 2 (the normal
// tail beyond |z|), z Phi(z) = max(z, 0) - q |z| for either sign of z.  erfc(x) for x >= 0 is
// Abramowitz and Stegun 7.1.26 (Hastings): t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) e^{-x^2} with
// t = 1 / (1 + p x), absolute error <= 1.5e-7.  Written in b = |z| sqrt(log2(e) / 2), so that
// e^{-x^2} = 2^{-b^2} and p x = p' b with p' = p / sqrt(log2 e); the coefficients carry the 1/2.
// The GELU so computed is within 2.2e-7 of the exact one (7e-8 |z|), far below BF16 resolution.
struct NormalTail {
  float q;  // erfc(|z| / sqrt 2) / 2
  float e;  // e^{-z^2 / 2}
};
__device__ __forceinline__ NormalTail normal_tail(float z) {
  const float b = fabsf(z) * 0.84932180028801907f;
  const float t = __builtin_amdgcn_rcpf(fmaf(0.27273748087922250f, b, 1.0f));
  const float e = __builtin_amdgcn_exp2f(-b * b);
  const float poly =
      t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 0.5307027145f, -0.7265760135f), 0.7107068705f), -0.142248368f), 0.127414796f);
  return NormalTail{poly * e, e};
}
__device__ __forceinline__ float gelu(float z) { return fmaf(-normal_tail(z).q, fabsf(z), fmaxf(z, 0.0f)); }
// d/dz z Phi(z) = Phi(z) + z phi(z), phi(z) = e^{-z^2/2} / sqrt(2 pi).
__device__ __forceinline__ float gelu_grad(float z) {
  const NormalTail nt = normal_tail(z);
  const float cdf = z >= 0.0f ? 1.0f - nt.q : nt.q;
  return fmaf(z * 0.39894228040143268f, nt.e, cdf);
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
                                                      int ldh, int kchunk) {
  __shared__ uint16_t sA[kTile][kSlice + 4];  // [m][k]
  __shared__ uint16_t sB[kTile][kSlice + 4];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  // split K: workgroup z sums k in [kb, ke) and writes its own M x ldo slab of outF
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  outF += (size_t)blockIdx.z * M * ldo;
  float16_t acc;
  for (int v = 0; v < 16; ++v) acc[v] = 0.0f;

  for (int k0 = kb; k0 < ke; k0 += kSlice) {
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
      if (gm < M && gk < ke) v = transA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
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
      if (gn < N && gk < ke) w = transB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
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

// out[i] = sum over s < parts of ws[s][i], in order (split-K partial products).
__global__ void split_reduce(const float *ws, int parts, int n, float *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float t = 0.0f;
  for (int q = 0; q < parts; ++q) t += ws[(size_t)q * n + i];
  out[i] = t;
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

// Partial column sums of dZ over a chunk of kSumRows rows: workgroup (x, y) sums 32 columns of
// rows [y*kSumRows, (y+1)*kSumRows) with 8 row groups (coalesced reads of 32 consecutive columns)
// and writes part[y][c]; a fixed-order reduction over chunks (split_reduce) gives db, so the result
// does not depend on scheduling.
constexpr int kSumRows = 64;
__global__ void column_sum(const float *dZ, int rows, int cols, float *part) {
  __shared__ float acc[8][32];
  const int cl = threadIdx.x & 31, rg = threadIdx.x >> 5;
  const int c = blockIdx.x * 32 + cl;
  const int r0 = blockIdx.y * kSumRows, r1 = min(rows, r0 + kSumRows);
  float s = 0.0f;
  if (c < cols)
    for (int r = r0 + rg; r < r1; r += 8) s += dZ[(size_t)r * cols + c];
  acc[rg][cl] = s;
  __syncthreads();
  if (rg == 0 && c < cols) {
    float t = 0.0f;
    for (int q = 0; q < 8; ++q) t += acc[q][cl];
    part[(size_t)blockIdx.y * cols + c] = t;
  }
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

// ---------------------------------------------------------------------------------------------
// Fused ensemble forward pass: one workgroup carries kFusedRows rows of the input through every
// layer of one network (blockIdx.y selects the network), the activations staying in LDS as BF16;
// v_mfma_f32_32x32x16_bf16 products with FP32 accumulation read A (activations) from LDS and B
// from a zero-padded BF16 weight copy in fragment order (K rounded up to kFusedK, N to 32): for
// column tile c and k-step s, the 64 lanes' 16-byte B fragments lie one after another, so a wave's
// fragment load reads 1 KB of consecutive memory.  Widths up to kFusedMax.
//
// Each of the kFusedWaves waves owns one 32-column tile of a layer (a layer of kFusedMax columns has
// exactly kFusedWaves tiles; narrower layers leave waves idle).  A wave's B fragments form one
// sequence over the layers it works in, read kFusedDepth k-steps ahead: the ring of fragments in
// flight runs across layer ends, so the first fragments of the next layer load during this layer's
// epilogue.  K is padded to a multiple of kFusedDepth k-steps, which keeps every ring slot a
// compile-time register.  Two workgroups share a CU (4 waves per SIMD), so one workgroup's
// epilogue (bias, GELU, BF16 conversion) overlaps the other's products.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

constexpr int kFusedRows = 32, kFusedMax = 256, kFusedPitch = kFusedMax + 8, kFusedLayers = 8, kFusedNets = 8;
constexpr int kFusedWaves = kFusedMax / 32, kFusedThreads = 64 * kFusedWaves;
constexpr int kFusedDepth = 4;             // k-steps of B fragments in flight per wave
constexpr int kFusedK = 16 * kFusedDepth;  // K padding of the packed weights

struct FusedLayer {
  const uint16_t *wt;  // [npad/32][k/16][64][8] BF16 fragments, k a multiple of kFusedK (pack_weights)
  const float *bias;   // [n]
  int k, n, npad;
};
struct FusedNet {
  FusedLayer layer[kFusedLayers];
  float *y;  // [rows][out] FP32
};
struct FusedArgs {
  FusedNet net[kFusedNets];
  const float *x;  // [rows][in] FP32, shared by all networks
  int rows, in, L;
};

// Barrier for LDS traffic only: the wave's LDS operations have completed, then the workgroup meets.
// (__syncthreads also waits for every global load in flight, which would drain the prefetched B
// fragments at each layer.)  The memory clobber keeps the compiler from moving LDS accesses across.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The products of one layer for one 32x32 output tile: STEPS k-steps of 16, A (activations) from LDS
// row `arow`, B from the ring of fragments in flight.  Each step refills its ring slot with the
// fragment kFusedDepth steps ahead: this layer's (`cur`) or, in the last steps, the next layer's
// (`nxt`).  Fully unrolled, so every slot and source is fixed at compile time.  A lane's fragments of
// consecutive k-steps are kFragStep elements apart.
constexpr int kFragStep = 64 * 8;
template <int STEPS>
__device__ __forceinline__ void layer_products(float16_t &acc, bf16x8_t (&ring)[kFusedDepth], const uint16_t *arow,
                                               const uint16_t *cur, const uint16_t *nxt) {
#pragma unroll
  for (int s = 0; s < STEPS; ++s) {
    const bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(arow + 16 * s);
    const bf16x8_t b = ring[s % kFusedDepth];
    const uint16_t *src = s + kFusedDepth < STEPS ? cur + kFragStep * (s + kFusedDepth)
                                                  : nxt + kFragStep * (s + kFusedDepth - STEPS);
    ring[s % kFusedDepth] = *reinterpret_cast<const bf16x8_t *>(src);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
}

__global__ __launch_bounds__(kFusedThreads) void mlp_forward_fused(FusedArgs args) {
  __shared__ __align__(16) uint16_t act[kFusedRows][kFusedPitch];
  const FusedNet &net = args.net[blockIdx.y];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r0 = blockIdx.x * kFusedRows;
  const int lr = lane & 31, lk = (lane >> 5) * 8;
  const int L = args.L;

  // This wave's B fragments: those of column tile `wave` of wt_l, in the layers where the wave owns
  // that tile.
  auto owns = [&](int l) { return wave * 32 < net.layer[l].npad; };
  auto base = [&](int l) {
    return net.layer[l].wt + ((size_t)wave * (net.layer[l].k / 16) * 64 + lane) * 8;
  };
  auto next_owned = [&](int l) {
    while (l < L && !owns(l)) ++l;
    return l;
  };
  bf16x8_t ring[kFusedDepth];
  int cl = next_owned(0);
  if (cl < L) {
    const uint16_t *p = base(cl);
#pragma unroll
    for (int d = 0; d < kFusedDepth; ++d) ring[d] = *reinterpret_cast<const bf16x8_t *>(p + kFragStep * d);
  }

  // inputs, BF16, zero outside the rows and the first `in` columns
  const int k0w = net.layer[0].k;
  for (int e = tid; e < kFusedRows * k0w; e += kFusedThreads) {
    const int m = e / k0w, k = e % k0w;
    float v = 0.0f;
    if (r0 + m < args.rows && k < args.in) v = args.x[(size_t)(r0 + m) * args.in + k];
    act[m][k] = f2bf(v);
  }
  lds_barrier();

  const uint16_t *arow = &act[lr][lk];
  for (int l = 0; l < L; ++l) {
    const FusedLayer ly = net.layer[l];
    const bool mine = l == cl;  // this wave owns column tile `wave` of this layer
    const bool last = l + 1 == L;
    const int col = wave * 32 + lr;
    float16_t acc;
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    float bz = 0.0f;
    if (mine) {
      if (col < ly.n) bz = ly.bias[col];  // loaded ahead of the ring's refills
      const int nl = next_owned(l + 1);
      const uint16_t *cur = base(l);
      const uint16_t *nxt = nl < L ? base(nl) : cur;  // past the last layer: harmless reloads
      switch (ly.k) {
        case 64: layer_products<4>(acc, ring, arow, cur, nxt); break;
        case 128: layer_products<8>(acc, ring, arow, cur, nxt); break;
        case 192: layer_products<12>(acc, ring, arow, cur, nxt); break;
        default: layer_products<16>(acc, ring, arow, cur, nxt); break;
      }
      cl = nl;
    }
    lds_barrier();  // every wave has read this layer's input
    if (mine) {
      if (last) {
        for (int v = 0; v < 16; ++v) {
          const int row = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          if (col < ly.n && r0 + row < args.rows) net.y[(size_t)(r0 + row) * ly.n + col] = acc[v] + bz;
        }
      } else {
        // columns n..npad-1 have zero weights and bias: GELU(0) = 0 pads the next layer's K
        for (int v = 0; v < 16; ++v) {
          const int row = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          act[row][col] = f2bf(gelu(acc[v] + bz));
        }
      }
    } else if (!last && col < net.layer[l + 1].k) {
      for (int v = 0; v < 16; ++v) act[8 * (v >> 2) + 4 * (lane >> 5) + (v & 3)][col] = 0;  // K padding
    }
    lds_barrier();
  }
}

// Fragment-order copy of the row-major BF16 weights W [K][N] for the fused pass (kpad a multiple of
// 16, npad of 32): element j of lane l's B fragment for column tile c and k-step s,
// wt[((c * kpad/16 + s) * 64 + l) * 8 + j], is W[16 s + 8 (l / 32) + j][32 c + l % 32], zero
// outside k < K, n < N.
__global__ void pack_weights(const uint16_t *W, int K, int N, uint16_t *wt, int kpad, int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad * kpad) return;
  const int j = i & 7, l = (i >> 3) & 63, f = i >> 9;
  const int s = f % (kpad / 16), c = f / (kpad / 16);
  const int n = 32 * c + (l & 31), k = 16 * s + 8 * (l >> 5) + j;
  wt[i] = (n < N && k < K) ? W[(size_t)k * N + n] : (uint16_t)0;
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
  // padded BF16 weights in fragment order for the fused forward pass; rebuilt when the weights change
  uint16_t *wt = nullptr;
  std::vector<size_t> wt_off;
  std::vector<int> k16, npad;
  bool wt_dirty = true;
  float *ws = nullptr;  // partial sums: split-K weight gradients, per-chunk bias gradients
  size_t ws_size = 0;   // floats
  float *xf = nullptr;  // fused-pass input staging [4096 x in]
  int xf_rows = 0;
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

// Largest number of split-K partial products of a weight gradient.
constexpr int kSplits = 8;

int max_width(const ymc_mlp *n) { return *std::max_element(n->widths.begin(), n->widths.end()); }

ymc_status ensure_capacity(ymc_mlp *n, int batch) {
  if (batch <= n->cap) return YMC_OK;
  {  // partial-sum workspace: kSplits weight-gradient slabs, or one bias-gradient row per row chunk
    const int L = (int)n->widths.size() - 1;
    size_t wmax = 0;
    for (int l = 0; l < L; ++l) wmax = std::max(wmax, (size_t)n->widths[l] * n->widths[l + 1]);
    const size_t need = std::max((size_t)kSplits * wmax, (size_t)((batch + kSumRows - 1) / kSumRows) * max_width(n));
    if (need > n->ws_size) {
      (void)hipFree(n->ws);
      n->ws = nullptr;
      n->ws_size = 0;
      ymc_status st;
      if ((st = dalloc(&n->ws, need))) return st;
      n->ws_size = need;
    }
  }
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
          int K, const float *bias, float *outF, int ldo, uint16_t *outH, int ldh, int splits = 1) {
  const int kchunk = splits > 1 ? ((K + splits - 1) / splits + kSlice - 1) / kSlice * kSlice : K;
  dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile, splits);
  if (epi == kStoreF32)
    hipLaunchKernelGGL(gemm_bf16<kStoreF32>, grid, dim3(kThreads), 0, s, A, lda, tA, B, ldb, tB, M, N, K, bias, outF,
                       ldo, outH, ldh, kchunk);
  else if (epi == kBiasGelu)
    hipLaunchKernelGGL(gemm_bf16<kBiasGelu>, grid, dim3(kThreads), 0, s, A, lda, tA, B, ldb, tB, M, N, K, bias, outF,
                       ldo, outH, ldh, kchunk);
  else
    hipLaunchKernelGGL(gemm_bf16<kBias>, grid, dim3(kThreads), 0, s, A, lda, tA, B, ldb, tB, M, N, K, bias, outF, ldo,
                       outH, ldh, kchunk);
}

// Weight gradient dW [M x N] = X^T dZ over `batch` rows: split over K (the batch) into up to
// kSplits partial products reduced in a fixed order when the output has few 64x64 tiles.
void weight_grad(ymc_mlp *n, const uint16_t *X, int M, const uint16_t *dZ, int N, int batch, float *dW);

void weight_grad(ymc_mlp *n, const uint16_t *X, int M, const uint16_t *dZ, int N, int batch, float *dW) {
  const int tiles = ((M + kTile - 1) / kTile) * ((N + kTile - 1) / kTile);
  int splits = std::min(kSplits, std::max(1, 128 / tiles));
  splits = std::min(splits, std::max(1, batch / (2 * kSlice)));
  if (splits <= 1) {
    gemm(n->stream, kStoreF32, X, M, 1, dZ, N, 0, M, N, batch, nullptr, dW, N, nullptr, 0);
    return;
  }
  gemm(n->stream, kStoreF32, X, M, 1, dZ, N, 0, M, N, batch, nullptr, n->ws, N, nullptr, 0, splits);
  const int cnt = M * N;
  hipLaunchKernelGGL(split_reduce, dim3((cnt + 255) / 256), dim3(256), 0, n->stream, n->ws, splits, cnt, dW);
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
    const int chunks = (batch + kSumRows - 1) / kSumRows;
    hipLaunchKernelGGL(column_sum, dim3((ow + 31) / 32, chunks), dim3(256), 0, n->stream, n->dZ, batch, ow, n->ws);
    hipLaunchKernelGGL(split_reduce, dim3((ow + 255) / 256), dim3(256), 0, n->stream, n->ws, chunks, ow,
                       n->g + n->b_off[l]);
    weight_grad(n, n->H[l], in, n->dZh, ow, batch, n->g + n->w_off[l]);
    if (l == 0) break;
    // dA_prev = dZ W^T, then dZ_prev = dA_prev * gelu'(Z_prev)
    gemm(n->stream, kStoreF32, n->dZh, ow, 0, n->ph + n->w_off[l], ow, 1, batch, in, ow, nullptr, n->dA, in, nullptr, 0);
    const int c2 = batch * in;
    hipLaunchKernelGGL(gelu_backward, dim3((c2 + 255) / 256), dim3(256), 0, n->stream, n->dA, n->Z[l - 1], n->dZ,
                       n->dZh, c2);
  }
  return blocks;
}

bool fused_ok(const ymc_mlp *n) {
  if ((int)n->widths.size() - 1 > kFusedLayers) return false;
  for (int w : n->widths)
    if (w > kFusedMax) return false;
  return true;
}

// Allocates the fragment-order weight copy of the fused pass (networks the fused pass serves;
// done when the network is created, so the passes themselves allocate nothing).
ymc_status alloc_wt(ymc_mlp *n) {
  if (n->wt || !fused_ok(n)) return YMC_OK;
  const int L = (int)n->widths.size() - 1;
  size_t total = 0;
  for (int l = 0; l < L; ++l) {
    n->k16.push_back((n->widths[l] + kFusedK - 1) / kFusedK * kFusedK);
    n->npad.push_back((n->widths[l + 1] + 31) / 32 * 32);
    n->wt_off.push_back(total);
    total += (size_t)n->k16[l] * n->npad[l];
  }
  return dalloc(&n->wt, total);
}

// Rebuilds the fragment-order weight copy from the BF16 weights if they changed.
ymc_status refresh_wt(ymc_mlp *n) {
  if (!n->wt_dirty) return YMC_OK;
  const int L = (int)n->widths.size() - 1;
  if (!n->wt) return fail(YMC_ERR_STATE, "fused pass: packed weights missing");
  for (int l = 0; l < L; ++l) {
    const int cnt = n->k16[l] * n->npad[l];
    hipLaunchKernelGGL(pack_weights, dim3((cnt + 255) / 256), dim3(256), 0, n->stream, n->ph + n->w_off[l],
                       n->widths[l], n->widths[l + 1], n->wt + n->wt_off[l], n->k16[l], n->npad[l]);
  }
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipStreamSynchronize(n->stream));
  n->wt_dirty = false;
  return YMC_OK;
}

// Arguments of the fused pass for `count` networks of identical widths.
FusedArgs fused_args(ymc_mlp *const *nets, int count, const float *x, int rows, float *const *y) {
  FusedArgs a{};
  const ymc_mlp *n0 = nets[0];
  a.L = (int)n0->widths.size() - 1;
  a.in = n0->widths[0];
  a.rows = rows;
  a.x = x;
  for (int q = 0; q < count; ++q) {
    const ymc_mlp *n = nets[q];
    for (int l = 0; l < a.L; ++l) {
      a.net[q].layer[l].wt = n->wt + n->wt_off[l];
      a.net[q].layer[l].bias = n->p + n->b_off[l];
      a.net[q].layer[l].k = n->k16[l];
      a.net[q].layer[l].n = n->widths[l + 1];
      a.net[q].layer[l].npad = n->npad[l];
    }
    a.net[q].y = y[q];
  }
  return a;
}

ymc_status check_ensemble(ymc_mlp *const *nets, int32_t count) {
  if (!nets || count < 1 || count > kFusedNets) return fail(YMC_ERR_ARG, "ensemble of 1..%d networks expected", kFusedNets);
  for (int q = 0; q < count; ++q) {
    if (!nets[q]) return fail(YMC_ERR_ARG, "null network %d", q);
    if (nets[q]->widths != nets[0]->widths || nets[q]->device != nets[0]->device)
      return fail(YMC_ERR_ARG, "ensemble networks must share widths and device");
  }
  if (!fused_ok(nets[0])) return fail(YMC_ERR_ARG, "fused pass needs widths <= %d and <= %d layers", kFusedMax, kFusedLayers);
  return YMC_OK;
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
      (st = dalloc(&n->v, n->n_params)) || (st = dalloc(&n->ph, n->n_params)) || (st = alloc_wt(n))) {
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
  if (n->wt) (void)hipFree(n->wt);
  if (n->ws) (void)hipFree(n->ws);
  if (n->xf) (void)hipFree(n->xf);
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
  n->wt_dirty = true;
  return YMC_OK;
}

ymc_status ymc_mlp_get_params(ymc_mlp *n, float *params, uint64_t count) {
  if (!n || !params || count != n->n_params) return fail(YMC_ERR_ARG, "ymc_mlp_get_params: expected %zu values", n ? n->n_params : 0);
  MLP_TRY(hipSetDevice(n->device));
  MLP_TRY(hipStreamSynchronize(n->stream));
  MLP_TRY(hipMemcpy(params, n->p, count * sizeof(float), hipMemcpyDeviceToHost));
  return YMC_OK;
}

// Fused pass over `rows` host rows for `count` networks, in launches of kFusedChunk rows.
constexpr int kFusedChunk = 4096;

// Fused pass through a staging buffer `stage` of chunk * (in + count * out) floats on the networks'
// device: each chunk of rows is copied in, run, and its outputs copied out.  Allocates nothing.
static ymc_status fused_forward_staged(ymc_mlp *const *nets, int32_t count, const float *x, int64_t rows, float *y,
                                       float *stage, int chunk) {
  ymc_mlp *n0 = nets[0];
  ymc_status st;
  for (int q = 0; q < count; ++q) {
    if ((st = refresh_wt(nets[q]))) return st;
    MLP_TRY(hipStreamSynchronize(nets[q]->stream));
  }
  const int L = (int)n0->widths.size() - 1;
  const int in = n0->widths[0], out = n0->widths[L];
  float *yd = stage + (size_t)chunk * in;
  for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
    const int b = (int)std::min<int64_t>(chunk, rows - r0);
    MLP_TRY(hipMemcpyAsync(stage, x + r0 * in, (size_t)b * in * sizeof(float), hipMemcpyHostToDevice, n0->stream));
    float *ys[kFusedNets];
    for (int q = 0; q < count; ++q) ys[q] = yd + (size_t)q * chunk * out;
    const FusedArgs a = fused_args(nets, count, stage, b, ys);
    hipLaunchKernelGGL(mlp_forward_fused, dim3((b + kFusedRows - 1) / kFusedRows, count), dim3(kFusedThreads), 0,
                       n0->stream, a);
    for (int q = 0; q < count; ++q)
      MLP_TRY(hipMemcpyAsync(y + ((size_t)q * rows + r0) * out, ys[q], (size_t)b * out * sizeof(float),
                             hipMemcpyDeviceToHost, n0->stream));
  }
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipStreamSynchronize(n0->stream));
  return YMC_OK;
}

static ymc_status fused_forward(ymc_mlp *const *nets, int32_t count, const float *x, int64_t rows, float *y) {
  ymc_mlp *n0 = nets[0];
  MLP_TRY(hipSetDevice(n0->device));
  ymc_status st;
  const int L = (int)n0->widths.size() - 1;
  const int in = n0->widths[0], out = n0->widths[L];
  if (n0->xf_rows < kFusedChunk) {
    if (n0->xf) (void)hipFree(n0->xf);
    n0->xf = nullptr;
    n0->xf_rows = 0;
    if ((st = dalloc(&n0->xf, (size_t)kFusedChunk * (in + kFusedNets * out)))) return st;
    n0->xf_rows = kFusedChunk;
  }
  return fused_forward_staged(nets, count, x, rows, y, n0->xf, kFusedChunk);
}

// Staging buffer of ymc_mlp_forward_ensemble_noalloc, one per device for the life of the process:
// the inputs and outputs of one chunk of rows.  Host code never uses the name `scratch` as an
// address; it uses the per-device address hipGetSymbolAddress returns (scratch_address).
constexpr size_t kScratchFloats = size_t(1) << 20;  // 4 MiB
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mutex;
static float *g_scratch_ptr[kScratchDevices];

// Address of `scratch` on `device`, which must be the current device; resolved once per device.
static ymc_status scratch_address(int device, float **out) {
  if (device < 0 || device >= kScratchDevices) return fail(YMC_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p)
      return fail(YMC_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(err));
    g_scratch_ptr[device] = static_cast<float *>(p);
  }
  *out = g_scratch_ptr[device];
  return YMC_OK;
}

ymc_status ymc_mlp_forward(ymc_mlp *n, const float *x, int64_t rows, float *y) {
  if (!n || !x || !y || rows < 0) return fail(YMC_ERR_ARG, "ymc_mlp_forward: bad arguments");
  if (fused_ok(n)) return fused_forward(&n, 1, x, rows, y);
  // layer-by-layer products for networks wider than the fused pass supports
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

ymc_status ymc_mlp_forward_ensemble(ymc_mlp *const *nets, int32_t count, const float *x, int64_t rows, float *y) {
  ymc_status st = check_ensemble(nets, count);
  if (st) return st;
  if (!x || !y || rows < 0) return fail(YMC_ERR_ARG, "ymc_mlp_forward_ensemble: bad arguments");
  return fused_forward(nets, count, x, rows, y);
}

ymc_status ymc_mlp_forward_ensemble_noalloc(ymc_mlp *const *nets, int32_t count, const float *x, int64_t rows,
                                            float *y) {
  ymc_status st = check_ensemble(nets, count);
  if (st) return st;
  if (!x || !y || rows < 0) return fail(YMC_ERR_ARG, "ymc_mlp_forward_ensemble_noalloc: bad arguments");
  ymc_mlp *n0 = nets[0];
  const int L = (int)n0->widths.size() - 1;
  const size_t per_row = (size_t)n0->widths[0] + (size_t)count * n0->widths[L];
  if (per_row * kFusedRows > kScratchFloats)
    return fail(YMC_ERR_ARG,
                "ymc_mlp_forward_ensemble_noalloc: inputs + count x outputs = %zu floats per row; the %zu-float "
                "staging buffer takes at most %zu",
                per_row, kScratchFloats, kScratchFloats / kFusedRows);
  const int chunk = (int)std::min<size_t>(kFusedChunk, kScratchFloats / per_row);
  MLP_TRY(hipSetDevice(n0->device));
  float *stage = nullptr;
  if ((st = scratch_address(n0->device, &stage))) return st;
  return fused_forward_staged(nets, count, x, rows, y, stage, chunk);
}

ymc_status ymc_mlp_time_forward(ymc_mlp *const *nets, int32_t count, int32_t rows, int32_t reps, double *ms_per_pass) {
  ymc_status st = check_ensemble(nets, count);
  if (st) return st;
  if (!ms_per_pass || rows < 1 || rows > kFusedChunk || reps < 1)
    return fail(YMC_ERR_ARG, "ymc_mlp_time_forward: 1..%d rows, reps >= 1", kFusedChunk);
  ymc_mlp *n0 = nets[0];
  // one chunk through the host path allocates the staging buffers and refreshes the weights
  const int L = (int)n0->widths.size() - 1;
  std::vector<float> xh((size_t)rows * n0->widths[0], 0.0f), yh((size_t)count * rows * n0->widths[L]);
  if ((st = fused_forward(nets, count, xh.data(), rows, yh.data()))) return st;
  float *ys[kFusedNets];
  for (int q = 0; q < count; ++q) ys[q] = n0->xf + (size_t)kFusedChunk * n0->widths[0] + (size_t)q * kFusedChunk * n0->widths[L];
  const FusedArgs a = fused_args(nets, count, n0->xf, rows, ys);
  const dim3 grid((rows + kFusedRows - 1) / kFusedRows, count);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MLP_TRY(hipEventCreate(&e0));
  MLP_TRY(hipEventCreate(&e1));
  MLP_TRY(hipEventRecord(e0, n0->stream));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(mlp_forward_fused, grid, dim3(kFusedThreads), 0, n0->stream, a);
  MLP_TRY(hipEventRecord(e1, n0->stream));
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
  n->wt_dirty = true;
  return YMC_OK;
}

}  // extern "C"
