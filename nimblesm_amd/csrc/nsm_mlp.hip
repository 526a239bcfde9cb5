// nsm_mlp.hip -- the surrogate network of the ensemble estimator: a fully
// connected network with GELU between layers, BF16 weights and activations,
// FP32 accumulation on gfx950's matrix cores (v_mfma_f32_16x16x32_bf16),
// FP32 master parameters trained by Adam.  Kernels and the nsm_mlp_* C ABI.
//
// Layer l (1..L) with width d_l: Z_l = H_{l-1} W_l^T + b_l (FP32), hidden
// layers H_l = bf16(gelu(Z_l)), output Y = Z_L.  H_0 = bf16(x).  Gradients of
// the mean squared error over batch and outputs run the same GEMM kernel:
// G_L = 2 (Y - T) / (B d_L); for l = L..1: D_l = bf16(G_l),
// dW_l = D_l^T H_{l-1}, db_l = sum_b G_l (FP32), and below the first layer
// G_{l-1} = (D_l W_l) * gelu'(Z_{l-1}).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/nimblesm_amd.h"

namespace nsm {
int set_last_error(int status, const char* msg);
}

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  f32x4 __attribute__((ext_vector_type(4)));

int fail(int status, const char* fmt, ...)
{
  char    buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return nsm::set_last_error(status, buf);
}

#define MLP_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t _e = (call);                                                                        \
    if (_e != hipSuccess) return fail(NSM_ERR_HIP, "%s: %s", #call, hipGetErrorString(_e));        \
  } while (0)

#define MLP_LAUNCHED()                                                                             \
  do {                                                                                             \
    hipError_t _e = hipGetLastError();                                                             \
    if (_e != hipSuccess) return fail(NSM_ERR_HIP, "kernel launch: %s", hipGetErrorString(_e));    \
  } while (0)

// GELU with the error function, and its derivative
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x)
{
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// ---------------------------------------------------------------------------
// GEMM: C[m][n] = sum_k A(m, k) B(k, n), A(m, k) = a[m a_m + k a_k],
// B(k, n) = b[k b_k + n b_n] (BF16), FP32 accumulation, with an epilogue.
// The network's products are small (M <= a few thousand rows, N, K <= 512),
// so a product's time is the latency of its loads, not its multiplies: a
// 256-thread workgroup computes a 32 x 32 tile of C (four waves of one
// 16 x 16 MFMA tile each) and stages up to 256 of K at once, every load of
// that range in flight together, so a product with K <= 256 waits for one
// round trip to memory instead of one per 32 of K.
// MFMA 16x16x32 operand layout: lane l holds row (A) / column (B) l % 16 and
// k = 8 (l / 16) .. 8 (l / 16) + 7; the result holds column l % 16, rows
// 4 (l / 16) .. 4 (l / 16) + 3.
// ---------------------------------------------------------------------------
constexpr int kTile = 32, kKStep = 32, kKStage = 256, kPad = 8, kGemmThreads = 256;
// chunks of 8 values one thread loads per operand and K stage
constexpr int kChunks = kTile * kKStage / 8 / kGemmThreads;
// weight-gradient products (K = batch) longer than this are split into at
// most kMaxSplits K ranges summed by k_split_sum
constexpr int kSplitAbove = 1024, kMaxSplits = 8;

enum Epilogue {
  kStoreC     = 0,  // C = v (FP32)
  kHidden     = 1,  // z = v + bias: C = z (FP32), H = bf16(gelu(z))
  kOutput     = 2,  // C = v + bias (FP32)
  kBackHidden = 3,  // g = v * gelu'(Z): G = g (FP32), D = bf16(g)
};

// How an operand tile is staged: 16-byte loads along k (contiguous in k),
// 16-byte loads along m / n written transposed into LDS (contiguous in m /
// n), or one element at a time.
enum Staging { kAlongK = 1, kAlongMN = 2, kScalar = 0 };

struct GemmArgs {
  int          M, N, K;
  const bf16*  a;
  int64_t      a_m, a_k;
  const bf16*  b;
  int64_t      b_k, b_n;
  float*       c;  // [M][ldc]; split-K: split z writes c + z * split_stride
  int64_t      ldc;
  const float* bias;  // [N]
  bf16*        h;     // [M][ldc]
  const float* z;     // [M][ldc]: pre-activation (kBackHidden)
  int          a_stage, b_stage;
  int          k_chunk;  // K range of one split (multiple of kKStep)
  int64_t      split_stride;
};

// One operand's share of a K stage: the 32 rows r0.. of the M or N side and
// k in [k0, k0 + 32 nk32) are cut into chunks of 8 values; chunk c of
// thread t is t + 256 c.  AlongK / scalar: chunk = (row, 8 k); AlongMN:
// chunk = (k, 8 rows).  Values past R or K are zero.
struct StageRegs {
  bf16x8 v[kChunks];
};

__device__ __forceinline__ void chunk_coords(int mode, int q, int& r, int& k)
{
  if (mode == kAlongMN) {
    k = q >> 2;
    r = (q & 3) * 8;
  } else {
    r = q & (kTile - 1);
    k = (q >> 5) * 8;
  }
}

__device__ __forceinline__ void stage_load(const bf16* p, int64_t s_r, int64_t s_k, int mode, int r0, int R, int k0,
                                           int K, int nchunks, StageRegs& g)
{
#pragma unroll
  for (int c = 0; c < kChunks; c++) {
    const int q = (int)threadIdx.x + c * kGemmThreads;
    if (q >= nchunks) continue;
    int rl, kl;
    chunk_coords(mode, q, rl, kl);
    const int r = r0 + rl, k = k0 + kl;
    if (mode == kAlongK) {
      if (r < R && k + 8 <= K) {
        g.v[c] = *reinterpret_cast<const bf16x8*>(p + (int64_t)r * s_r + k);
        continue;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) g.v[c][j] = (r < R && k + j < K) ? p[(int64_t)r * s_r + (k + j)] : (bf16)0.0f;
    } else if (mode == kAlongMN) {
      if (k < K && r + 8 <= R) {
        g.v[c] = *reinterpret_cast<const bf16x8*>(p + (int64_t)k * s_k + r);
        continue;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) g.v[c][j] = (k < K && r + j < R) ? p[(int64_t)k * s_k + (r + j)] : (bf16)0.0f;
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++)
        g.v[c][j] = (r < R && k + j < K) ? p[(int64_t)r * s_r + (int64_t)(k + j) * s_k] : (bf16)0.0f;
    }
  }
}

template <int kLd>
__device__ __forceinline__ void stage_store(const StageRegs& g, int mode, int nchunks, bf16 (*S)[kLd])
{
#pragma unroll
  for (int c = 0; c < kChunks; c++) {
    const int q = (int)threadIdx.x + c * kGemmThreads;
    if (q >= nchunks) continue;
    int r, k;
    chunk_coords(mode, q, r, k);
    if (mode == kAlongMN) {
#pragma unroll
      for (int j = 0; j < 8; j++) S[r + j][k] = g.v[c][j];
    } else {
      *reinterpret_cast<bf16x8*>(&S[r][k]) = g.v[c];
    }
  }
}

// Tile (bx, by) of C (columns 32 bx.., rows 32 by..) over split bz of the
// K range, by the whole workgroup (bx, by, bz uniform).
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, int mode, int bx, int by, int bz)
{
  constexpr int kLd = kKStage + kPad;
  __shared__ __attribute__((aligned(16))) bf16 As[kTile][kLd];
  __shared__ __attribute__((aligned(16))) bf16 Bs[kTile][kLd];
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const int m0 = by * kTile, n0 = bx * kTile;
  const int wm = wave >> 1, wn = wave & 1;
  const int kb = bz * g.k_chunk, ke = min(g.K, kb + g.k_chunk);
  f32x4     acc = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kb; k0 < ke; k0 += kKStage) {
    const int nk32    = (min(kKStage, ke - k0) + kKStep - 1) / kKStep;  // 32-wide steps of this stage
    const int nchunks = kTile * nk32 * kKStep / 8;
    StageRegs ra, rb;
    stage_load(g.a, g.a_m, g.a_k, g.a_stage, m0, g.M, k0, ke, nchunks, ra);
    stage_load(g.b, g.b_n, g.b_k, g.b_stage, n0, g.N, k0, ke, nchunks, rb);
    if (k0 > kb) __syncthreads();  // the previous stage's LDS reads are done
    stage_store<kLd>(ra, g.a_stage, nchunks, As);
    stage_store<kLd>(rb, g.b_stage, nchunks, Bs);
    __syncthreads();
    for (int s = 0; s < nk32; s++) {
      const bf16x8 af  = *reinterpret_cast<const bf16x8*>(&As[wm * 16 + (lane & 15)][s * kKStep + 8 * (lane >> 4)]);
      const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(&Bs[wn * 16 + (lane & 15)][s * kKStep + 8 * (lane >> 4)]);
      acc              = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr, acc, 0, 0, 0);
    }
  }

  float* const c = g.c + (int64_t)bz * g.split_stride;
  const int    n = n0 + wn * 16 + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int m = m0 + wm * 16 + 4 * (lane >> 4) + r;
    if (m >= g.M || n >= g.N) continue;
    const int64_t o = (int64_t)m * g.ldc + n;
    float         v = acc[r];
    if (mode == kStoreC) {
      c[o] = v;
    } else if (mode == kHidden) {
      v += g.bias[n];
      c[o]   = v;
      g.h[o] = (bf16)gelu(v);
    } else if (mode == kOutput) {
      c[o] = v + g.bias[n];
    } else {  // kBackHidden
      const float gr = v * gelu_grad(g.z[o]);
      c[o]           = gr;
      g.h[o]         = (bf16)gr;
    }
  }
}

// Row tiles run along the grid's x dimension (up to 2^31 - 1 workgroups; y
// and z are limited to 65535), column tiles along y, K ranges along z.
__global__ __launch_bounds__(kGemmThreads) void k_gemm_bf16(GemmArgs g, int mode)
{
  gemm_tile(g, mode, (int)blockIdx.y, (int)blockIdx.x, (int)blockIdx.z);
}

// db[n] = sum over the batch of G[b][n] for the four columns 4 w .. 4 w + 3
// of workgroup w: one wave per column, lane l sums rows l, l + 64, ... in
// ascending order, then a fixed shuffle tree
__device__ __forceinline__ void bias_grad_columns(const float* __restrict__ G, int64_t B, int N,
                                                  float* __restrict__ db, int w)
{
  const int n = w * 4 + (int)(threadIdx.x >> 6);
  if (n >= N) return;
  const int lane = (int)threadIdx.x & 63;
  float     s    = 0.0f;
  for (int64_t b = lane; b < B; b += 64) s += G[b * N + n];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_down(s, m, 64);
  if (lane == 0) db[n] = s;
}

// The independent parts of one layer's backward pass in one launch:
// workgroups [0, tiles[0]) compute tiles of product 0 (the weight
// gradient), the next tiles[1] tiles of product 1 (the input gradient,
// absent below the first layer), the rest the bias gradient.  Products are
// unsplit (K = batch <= kSplitAbove).
struct LayerBackward {
  GemmArgs     prod[2];
  int          mode[2];
  int          tiles[2], tiles_x[2];
  const float* G;  // [B][N] FP32 gradient of the layer's output
  int64_t      B;
  int          N;
  float*       db;
};

__global__ __launch_bounds__(kGemmThreads) void k_layer_backward(LayerBackward lb)
{
  int w = (int)blockIdx.x;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    if (w < lb.tiles[q]) {  // uniform across the workgroup
      gemm_tile(lb.prod[q], lb.mode[q], w % lb.tiles_x[q], w / lb.tiles_x[q], 0);
      return;
    }
    w -= lb.tiles[q];
  }
  bias_grad_columns(lb.G, lb.B, lb.N, lb.db, w);
}

// out[i] = sum over splits s of part[s * stride + i], ascending s
__global__ __launch_bounds__(256) void k_split_sum(const float* __restrict__ part, int splits, int64_t stride,
                                                   int64_t n, float* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int z = 0; z < splits; z++) s += part[(int64_t)z * stride + i];
  out[i] = s;
}

// FP32 -> BF16 (round to nearest even), n values
__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ x, bf16* __restrict__ y, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (bf16)x[i];
}

// Gradient G = 2 (y - t) / n of the mean squared error over n outputs (also
// as BF16), and per workgroup the sum of squared errors (fixed tree).
constexpr int kMseBlocks = 64;
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t, int64_t n,
                                             float* __restrict__ grad, bf16* __restrict__ grad_bf,
                                             double* __restrict__ partial)
{
  __shared__ double red[256];
  double            s     = 0.0;
  const float       scale = 2.0f / (float)n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float e = y[i] - t[i];
    s += (double)e * (double)e;
    const float gr = scale * e;
    grad[i]        = gr;
    grad_bf[i]     = (bf16)gr;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// loss of pass p = sum of its kMseBlocks partials (ascending) / n, for
// passes 0 .. gridDim.x - 1 (partial[p kMseBlocks + b])
__global__ __launch_bounds__(64) void k_mse_total(const double* __restrict__ partial, int64_t n,
                                                  double* __restrict__ loss_out)
{
  if (threadIdx.x != 0) return;
  const double* p = partial + (int64_t)blockIdx.x * kMseBlocks;
  double        s = 0.0;
  for (int b = 0; b < kMseBlocks; b++) s += p[b];
  loss_out[blockIdx.x] = s / (double)n;
}

// bias gradient alone (batches longer than kSplitAbove)
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ G, int64_t B, int N,
                                                   float* __restrict__ db)
{
  bias_grad_columns(G, B, N, db, (int)blockIdx.x);
}

// Adam on the FP32 master parameters; BF16 copies for the forward pass.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m1, float* __restrict__ m2,
                                              bf16* __restrict__ p_bf, int64_t n, float lr, float b1, float b2,
                                              float c1, float c2, float eps)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a  = b1 * m1[i] + (1.0f - b1) * gi;
  const float v  = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i]          = a;
  m2[i]          = v;
  const float pi = p[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  p[i]           = pi;
  p_bf[i]        = (bf16)pi;
}

inline unsigned blocks_of(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct nsm_mlp {
  int                  device = 0;
  hipStream_t          stream = nullptr;
  std::vector<int>     width;              // d_0 .. d_L
  std::vector<int64_t> w_off, b_off;       // offsets of W_l, b_l in the flat parameter array
  int64_t              num_params = 0;
  float*               p   = nullptr;      // FP32 master parameters
  bf16*                pb  = nullptr;      // BF16 copy
  float*               g   = nullptr;      // gradients
  float *              m1 = nullptr, *m2 = nullptr;
  int64_t              adam_t = 0;
  // activations for a batch of `cap` rows
  int64_t              cap = 0;
  std::vector<float*>  z;   // Z_l, l = 1..L   [cap][d_l] (index l-1)
  std::vector<bf16*>   h;   // H_l, l = 0..L-1 [cap][d_l]
  float*               gf = nullptr;  // G (FP32) [cap][max d]
  bf16*                gb = nullptr;  // D (BF16) [cap][max d]
  float*               gf2 = nullptr;
  bf16*                gb2 = nullptr;
  float*               x32 = nullptr;  // input [cap][d_0]
  float*               t32 = nullptr;  // targets [cap][d_L]
  double*              loss = nullptr; // [loss_cap]
  int64_t              loss_cap = 0;
  double*              loss_part = nullptr;  // [min(loss_cap, kLossChunk)][kMseBlocks]
  float*               split = nullptr;      // split-K partial products
  int64_t              split_cap = 0;

  int L() const { return (int)width.size() - 1; }
  int max_width() const { return *std::max_element(width.begin(), width.end()); }
};

namespace {

// staging mode of an operand with rows at stride s_r and k at stride s_k
int staging(const bf16* p, int64_t s_r, int64_t s_k)
{
  const bool aligned = ((uintptr_t)p % 16) == 0;
  if (aligned && s_k == 1 && s_r % 8 == 0) return kAlongK;
  if (aligned && s_r == 1 && s_k % 8 == 0) return kAlongMN;
  return kScalar;
}

// Staging modes and K ranges of a product cut into `splits` K ranges
// (clamped to its 32-wide steps); returns the split count.
int prepare(GemmArgs& a, int splits)
{
  a.a_stage = staging(a.a, a.a_m, a.a_k);
  a.b_stage = staging(a.b, a.b_n, a.b_k);
  splits    = std::max(1, std::min(splits, (a.K + kKStep - 1) / kKStep));
  a.k_chunk = ((a.K + splits - 1) / splits + kKStep - 1) / kKStep * kKStep;
  return splits;
}

// C = A B with the epilogue `mode`.  splits > 1 (kStoreC only): the K range
// is cut into `splits` parts written to net->split (one M x N block each)
// and summed in ascending order into a.c.
int mlp_gemm(nsm_mlp* net, GemmArgs a, int mode, int splits = 1)
{
  splits         = prepare(a, splits);
  float* final_c = a.c;
  const int64_t mn = (int64_t)a.M * a.ldc;
  if (splits > 1) {
    if (mode != kStoreC || mn * splits > net->split_cap) return fail(NSM_ERR_STATE, "split-K scratch too small");
    a.c            = net->split;
    a.split_stride = mn;
  }
  const dim3 grid(blocks_of(a.M, kTile), blocks_of(a.N, kTile), (unsigned)splits);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(kGemmThreads), 0, net->stream, a, mode);
  MLP_LAUNCHED();
  if (splits > 1) {
    hipLaunchKernelGGL(k_split_sum, dim3(blocks_of(mn, 256)), dim3(256), 0, net->stream, net->split, splits, mn, mn,
                       final_c);
    MLP_LAUNCHED();
  }
  return NSM_OK;
}

void free_batch(nsm_mlp* net)
{
  for (auto* q : net->z) (void)hipFree(q);
  for (auto* q : net->h) (void)hipFree(q);
  net->z.clear();
  net->h.clear();
  for (void* q : {(void*)net->gf, (void*)net->gb, (void*)net->gf2, (void*)net->gb2, (void*)net->x32,
                  (void*)net->t32})
    if (q) (void)hipFree(q);
  net->gf = nullptr;
  net->gb = nullptr;
  net->gf2 = nullptr;
  net->gb2 = nullptr;
  net->x32 = nullptr;
  net->t32 = nullptr;
  net->cap = 0;
}

int ensure_batch(nsm_mlp* net, int64_t B)
{
  if (B <= net->cap) return NSM_OK;
  free_batch(net);
  const int     L  = net->L();
  const int64_t mw = net->max_width();
  net->z.assign((size_t)L, nullptr);
  net->h.assign((size_t)L, nullptr);
  for (int l = 0; l < L; l++) {
    MLP_HIP(hipMalloc((void**)&net->z[(size_t)l], (size_t)(B * net->width[(size_t)l + 1]) * sizeof(float)));
    MLP_HIP(hipMalloc((void**)&net->h[(size_t)l], (size_t)(B * net->width[(size_t)l]) * sizeof(bf16)));
  }
  MLP_HIP(hipMalloc((void**)&net->gf, (size_t)(B * mw) * sizeof(float)));
  MLP_HIP(hipMalloc((void**)&net->gb, (size_t)(B * mw) * sizeof(bf16)));
  MLP_HIP(hipMalloc((void**)&net->gf2, (size_t)(B * mw) * sizeof(float)));
  MLP_HIP(hipMalloc((void**)&net->gb2, (size_t)(B * mw) * sizeof(bf16)));
  MLP_HIP(hipMalloc((void**)&net->x32, (size_t)(B * net->width[0]) * sizeof(float)));
  MLP_HIP(hipMalloc((void**)&net->t32, (size_t)(B * net->width[(size_t)L]) * sizeof(float)));
  net->cap = B;
  return NSM_OK;
}

// Forward pass of B rows already in x32 (converted to BF16 into h[0] unless
// convert_input is false: training converts once); output Z_L in z[L-1].
int forward(nsm_mlp* net, int64_t B, bool convert_input = true)
{
  const int L  = net->L();
  const int d0 = net->width[0];
  if (convert_input) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_of(B * d0, 256)), dim3(256), 0, net->stream, net->x32, net->h[0],
                       B * d0);
    MLP_LAUNCHED();
  }
  for (int l = 1; l <= L; l++) {
    const int din = net->width[(size_t)l - 1], dout = net->width[(size_t)l];
    GemmArgs  a{};
    a.M    = (int)B;
    a.N    = dout;
    a.K    = din;
    a.a    = net->h[(size_t)l - 1];
    a.a_m  = din;
    a.a_k  = 1;
    a.b    = net->pb + net->w_off[(size_t)l - 1];  // B(k, n) = W_l[n][k]
    a.b_k  = 1;
    a.b_n  = din;
    a.c    = net->z[(size_t)l - 1];
    a.ldc  = dout;
    a.bias = net->p + net->b_off[(size_t)l - 1];
    a.h    = l < L ? net->h[(size_t)l] : nullptr;
    int st = mlp_gemm(net, a, l < L ? kHidden : kOutput);
    if (st) return st;
  }
  return NSM_OK;
}

// Gradients of every parameter into net->g from the forward pass against
// t32; the pass's kMseBlocks squared-error partials go to `partial`
// (device; k_mse_total turns them into the loss).  With at most
// kSplitAbove batch rows each layer is one launch (k_layer_backward: weight
// gradient, input gradient and bias gradient side by side), else one per
// product.
int backward(nsm_mlp* net, int64_t B, double* partial)
{
  const int L  = net->L();
  const int dL = net->width[(size_t)L];
  hipLaunchKernelGGL(k_mse, dim3(kMseBlocks), dim3(256), 0, net->stream, net->z[(size_t)L - 1], net->t32, B * dL,
                     net->gf, net->gb, partial);
  MLP_LAUNCHED();
  float* gcur = net->gf;
  bf16*  dcur = net->gb;
  float* gnext = net->gf2;
  bf16*  dnext = net->gb2;
  for (int l = L; l >= 1; l--) {
    const int din = net->width[(size_t)l - 1], dout = net->width[(size_t)l];
    // dW_l[n][k] = sum_b D_l[b][n] H_{l-1}[b][k]
    GemmArgs a{};
    a.M   = dout;
    a.N   = din;
    a.K   = (int)B;
    a.a   = dcur;
    a.a_m = 1;
    a.a_k = dout;
    a.b   = net->h[(size_t)l - 1];
    a.b_k = din;
    a.b_n = 1;
    a.c   = net->g + net->w_off[(size_t)l - 1];
    a.ldc = din;
    // G_{l-1}[b][k] = (sum_n D_l[b][n] W_l[n][k]) gelu'(Z_{l-1}[b][k])
    GemmArgs d{};
    d.M   = (int)B;
    d.N   = din;
    d.K   = dout;
    d.a   = dcur;
    d.a_m = dout;
    d.a_k = 1;
    d.b   = net->pb + net->w_off[(size_t)l - 1];
    d.b_k = din;
    d.b_n = 1;
    d.c   = gnext;
    d.ldc = din;
    d.h   = dnext;
    d.z   = l > 1 ? net->z[(size_t)l - 2] : nullptr;
    float* const db = net->g + net->b_off[(size_t)l - 1];
    if (B <= kSplitAbove) {
      LayerBackward lb{};
      lb.prod[0]    = a;
      lb.mode[0]    = kStoreC;
      prepare(lb.prod[0], 1);
      lb.tiles_x[0] = (int)blocks_of(a.N, kTile);
      lb.tiles[0]   = lb.tiles_x[0] * (int)blocks_of(a.M, kTile);
      if (l > 1) {
        lb.prod[1]    = d;
        lb.mode[1]    = kBackHidden;
        prepare(lb.prod[1], 1);
        lb.tiles_x[1] = (int)blocks_of(d.N, kTile);
        lb.tiles[1]   = lb.tiles_x[1] * (int)blocks_of(d.M, kTile);
      }
      lb.G  = gcur;
      lb.B  = B;
      lb.N  = dout;
      lb.db = db;
      const unsigned grid = (unsigned)(lb.tiles[0] + lb.tiles[1]) + blocks_of(dout, 4);
      hipLaunchKernelGGL(k_layer_backward, dim3(grid), dim3(kGemmThreads), 0, net->stream, lb);
      MLP_LAUNCHED();
    } else {
      const int splits = (int)std::min<int64_t>(kMaxSplits, (B + kSplitAbove - 1) / kSplitAbove);
      int       st     = mlp_gemm(net, a, kStoreC, splits);
      if (st) return st;
      hipLaunchKernelGGL(k_bias_grad, dim3(blocks_of(dout, 4)), dim3(256), 0, net->stream, gcur, B, dout, db);
      MLP_LAUNCHED();
      if (l > 1 && (st = mlp_gemm(net, d, kBackHidden))) return st;
    }
    std::swap(gcur, gnext);
    std::swap(dcur, dnext);
  }
  return NSM_OK;
}

uint64_t splitmix(uint64_t& s)
{
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z          = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int upload_params(nsm_mlp* net, const float* host)
{
  MLP_HIP(hipMemcpyAsync(net->p, host, (size_t)net->num_params * sizeof(float), hipMemcpyHostToDevice,
                         net->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_of(net->num_params, 256)), dim3(256), 0, net->stream, net->p, net->pb,
                     net->num_params);
  MLP_LAUNCHED();
  MLP_HIP(hipMemsetAsync(net->m1, 0, (size_t)net->num_params * sizeof(float), net->stream));
  MLP_HIP(hipMemsetAsync(net->m2, 0, (size_t)net->num_params * sizeof(float), net->stream));
  net->adam_t = 0;
  MLP_HIP(hipStreamSynchronize(net->stream));
  return NSM_OK;
}

int load_rows(nsm_mlp* net, int64_t B, const float* x, const float* t)
{
  int st = ensure_batch(net, B);
  if (st) return st;
  MLP_HIP(hipMemcpyAsync(net->x32, x, (size_t)(B * net->width[0]) * sizeof(float), hipMemcpyHostToDevice,
                         net->stream));
  if (t)
    MLP_HIP(hipMemcpyAsync(net->t32, t, (size_t)(B * net->width.back()) * sizeof(float), hipMemcpyHostToDevice,
                           net->stream));
  return NSM_OK;
}

// squared-error partials of up to kLossChunk passes are kept on the device
// and turned into losses by one k_mse_total launch per chunk
constexpr int kLossChunk = 4096;

// losses for `passes` forward/backward passes and partials for
// min(passes, kLossChunk) of them
int ensure_loss(nsm_mlp* net, int64_t passes)
{
  if (passes <= net->loss_cap) return NSM_OK;
  if (net->loss) MLP_HIP(hipFree(net->loss));
  if (net->loss_part) MLP_HIP(hipFree(net->loss_part));
  net->loss      = nullptr;
  net->loss_part = nullptr;
  net->loss_cap  = 0;
  MLP_HIP(hipMalloc((void**)&net->loss, (size_t)passes * sizeof(double)));
  MLP_HIP(hipMalloc((void**)&net->loss_part,
                    (size_t)(std::min<int64_t>(passes, kLossChunk) * kMseBlocks) * sizeof(double)));
  net->loss_cap = passes;
  return NSM_OK;
}

}  // namespace

extern "C" {

int nsm_mlp_create(int num_layers, const int* widths, uint64_t seed, nsm_mlp** out)
{
  if (!out || !widths) return fail(NSM_ERR_ARG, "null argument");
  *out = nullptr;
  if (num_layers < 1 || num_layers > 16) return fail(NSM_ERR_ARG, "1 to 16 layers");
  for (int l = 0; l <= num_layers; l++)
    if (widths[l] < 1 || widths[l] > 4096) return fail(NSM_ERR_ARG, "layer widths must be in [1, 4096]");
  nsm_mlp* net = new nsm_mlp();
  net->width.assign(widths, widths + num_layers + 1);
  int64_t off = 0;
  for (int l = 1; l <= num_layers; l++) {
    net->w_off.push_back(off);
    off += (int64_t)widths[l] * widths[l - 1];
    net->b_off.push_back(off);
    off += widths[l];
  }
  net->num_params = off;
  auto cleanup    = [&](int st) {
    nsm_mlp_destroy(net);
    return st;
  };
  if (hipGetDevice(&net->device) != hipSuccess) return cleanup(fail(NSM_ERR_HIP, "no HIP device"));
  if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(NSM_ERR_HIP, "hipStreamCreate failed"));
  const size_t n = (size_t)net->num_params;
  int64_t max_w = 0;
  for (int l = 1; l <= num_layers; l++) max_w = std::max<int64_t>(max_w, (int64_t)widths[l] * widths[l - 1]);
  net->split_cap = kMaxSplits * max_w;
  if (hipMalloc((void**)&net->p, n * 4) != hipSuccess || hipMalloc((void**)&net->pb, n * 2) != hipSuccess ||
      hipMalloc((void**)&net->g, n * 4) != hipSuccess || hipMalloc((void**)&net->m1, n * 4) != hipSuccess ||
      hipMalloc((void**)&net->m2, n * 4) != hipSuccess ||
      hipMalloc((void**)&net->split, (size_t)net->split_cap * 4) != hipSuccess)
    return cleanup(fail(NSM_ERR_NOMEM, "device allocation of the parameters failed"));
  // uniform(-a, a) weights with a = sqrt(6 / (fan_in + fan_out)), zero biases
  std::vector<float> init(n, 0.0f);
  uint64_t           s = seed;
  for (int l = 1; l <= num_layers; l++) {
    const double a = std::sqrt(6.0 / (double)(widths[l - 1] + widths[l]));
    for (int64_t i = 0; i < (int64_t)widths[l] * widths[l - 1]; i++) {
      const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      init[(size_t)(net->w_off[(size_t)l - 1] + i)] = (float)((2.0 * u - 1.0) * a);
    }
  }
  int st = upload_params(net, init.data());
  if (st) return cleanup(st);
  *out = net;
  return NSM_OK;
}

int nsm_mlp_destroy(nsm_mlp* net)
{
  if (!net) return NSM_OK;
  (void)hipSetDevice(net->device);
  if (net->stream) (void)hipStreamSynchronize(net->stream);
  free_batch(net);
  for (void* q : {(void*)net->p, (void*)net->pb, (void*)net->g, (void*)net->m1, (void*)net->m2, (void*)net->loss,
                  (void*)net->loss_part, (void*)net->split})
    if (q) (void)hipFree(q);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
  return NSM_OK;
}

int nsm_mlp_num_parameters(nsm_mlp* net, int64_t* n)
{
  if (!net || !n) return fail(NSM_ERR_ARG, "null argument");
  *n = net->num_params;
  return NSM_OK;
}

int nsm_mlp_get_parameters(nsm_mlp* net, float* out)
{
  if (!net || !out) return fail(NSM_ERR_ARG, "null argument");
  MLP_HIP(hipSetDevice(net->device));
  MLP_HIP(hipMemcpyAsync(out, net->p, (size_t)net->num_params * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  MLP_HIP(hipStreamSynchronize(net->stream));
  return NSM_OK;
}

int nsm_mlp_set_parameters(nsm_mlp* net, const float* in)
{
  if (!net || !in) return fail(NSM_ERR_ARG, "null argument");
  MLP_HIP(hipSetDevice(net->device));
  return upload_params(net, in);
}

int nsm_mlp_forward(nsm_mlp* net, int64_t batch, const float* x, float* y)
{
  if (!net || (batch > 0 && (!x || !y))) return fail(NSM_ERR_ARG, "null argument");
  if (batch < 0 || batch > (1 << 24)) return fail(NSM_ERR_ARG, "batch out of range");
  if (batch == 0) return NSM_OK;
  MLP_HIP(hipSetDevice(net->device));
  int st = load_rows(net, batch, x, nullptr);
  if (st || (st = forward(net, batch))) return st;
  MLP_HIP(hipMemcpyAsync(y, net->z.back(), (size_t)(batch * net->width.back()) * sizeof(float),
                         hipMemcpyDeviceToHost, net->stream));
  MLP_HIP(hipStreamSynchronize(net->stream));
  return NSM_OK;
}

int nsm_mlp_loss_gradients(nsm_mlp* net, int64_t batch, const float* x, const float* t, double* loss, float* grad)
{
  if (!net || !x || !t || !loss || !grad) return fail(NSM_ERR_ARG, "null argument");
  if (batch < 1 || batch > (1 << 24)) return fail(NSM_ERR_ARG, "batch out of range");
  MLP_HIP(hipSetDevice(net->device));
  int st = load_rows(net, batch, x, t);
  if (st) return st;
  const int64_t n = batch * net->width.back();
  if ((st = ensure_loss(net, 1)) || (st = forward(net, batch)) || (st = backward(net, batch, net->loss_part)))
    return st;
  hipLaunchKernelGGL(k_mse_total, dim3(1), dim3(64), 0, net->stream, net->loss_part, n, net->loss);
  MLP_LAUNCHED();
  MLP_HIP(hipMemcpyAsync(loss, net->loss, sizeof(double), hipMemcpyDeviceToHost, net->stream));
  MLP_HIP(hipMemcpyAsync(grad, net->g, (size_t)net->num_params * sizeof(float), hipMemcpyDeviceToHost,
                         net->stream));
  MLP_HIP(hipStreamSynchronize(net->stream));
  return NSM_OK;
}

int nsm_mlp_train(nsm_mlp* net, int64_t batch, const float* x, const float* t, int iterations, double lr,
                  double lr_final, double* loss_history)
{
  if (!net || !x || !t) return fail(NSM_ERR_ARG, "null argument");
  if (batch < 1 || batch > (1 << 24) || iterations < 0) return fail(NSM_ERR_ARG, "invalid batch or iterations");
  if (!(lr > 0.0) || !(lr_final > 0.0)) return fail(NSM_ERR_ARG, "learning rates must be positive");
  MLP_HIP(hipSetDevice(net->device));
  int st = load_rows(net, batch, x, t);
  if (st) return st;
  if ((st = ensure_loss(net, iterations))) return st;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < iterations; it++) {
    // cosine decay from lr to lr_final
    const double c   = 0.5 * (1.0 + std::cos(M_PI * (double)it / (double)std::max(iterations - 1, 1)));
    const float  lri = (float)(lr_final + (lr - lr_final) * c);
    net->adam_t++;
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)net->adam_t));
    const float c2 = (float)(1.0 - std::pow((double)b2, (double)net->adam_t));
    if ((st = forward(net, batch, it == 0)) ||
        (st = backward(net, batch, net->loss_part + (int64_t)(it % kLossChunk) * kMseBlocks)))
      return st;
    hipLaunchKernelGGL(k_adam, dim3(blocks_of(net->num_params, 256)), dim3(256), 0, net->stream, net->p, net->g,
                       net->m1, net->m2, net->pb, net->num_params, lri, b1, b2, c1, c2, eps);
    MLP_LAUNCHED();
    // the losses of a full chunk of iterations (or of the last one)
    if ((it + 1) % kLossChunk == 0 || it + 1 == iterations) {
      const int first = it - it % kLossChunk;
      hipLaunchKernelGGL(k_mse_total, dim3((unsigned)(it + 1 - first)), dim3(64), 0, net->stream, net->loss_part,
                         batch * net->width.back(), net->loss + first);
      MLP_LAUNCHED();
    }
  }
  if (loss_history && iterations > 0)
    MLP_HIP(hipMemcpyAsync(loss_history, net->loss, (size_t)iterations * sizeof(double), hipMemcpyDeviceToHost,
                           net->stream));
  MLP_HIP(hipStreamSynchronize(net->stream));
  return NSM_OK;
}

}  // extern "C"
