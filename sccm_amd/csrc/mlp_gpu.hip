// Learned surrogate: a fully connected network (SiLU hidden layers, linear
// output, FP32 parameters) with hand-written HIP kernels on the MFMA units
// for the forward pass, backpropagation and a bfloat16 inference path, plus
// the loss and Adam.  C-ABI in include/pnamp_mlp.h.
//
// Per layer l with inputs a (B x n_in) and outputs z = a W^T + b (B x n_out):
//   k_gemm_f32<EPI>  C = op(A) op(B) on v_mfma_f32_32x32x2_f32, a 64 x 64 tile
//                    of C per workgroup (a 32 x 32 tile per wave), K in
//                    chunks of 16 staged through LDS; epilogues:
//                      EPI_FWD   z = C + b, a' = silu(z) (identity on the last layer)
//                      EPI_DATA  d_prev = C * silu'(z_prev)       (C = d W)
//                      EPI_STORE C as is (split-K partials of dW = d^T a_prev)
//   k_splitk_sum     dW = sum of the partials, db = column sums of d
//   k_loss           d = 2 w_o (f - y) / (B n_out) and block sums of the loss
//   k_adam           the Adam update over the flat parameter vector
//   k_mlp_infer_bf16 the whole network for 64 samples per workgroup on
//                    v_mfma_f32_32x32x16_bf16: activations in LDS as bfloat16,
//                    weights in bfloat16 fragment order (k_pack_bf16), FP32
//                    accumulation, bias and SiLU in FP32.
// Every sum has a fixed order (no atomics): results are reproducible.
// silu(z) = z s(z), silu'(z) = s(z) (1 + z (1 - s(z))), s = logistic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/pnamp_mlp.h"

namespace {

constexpr int PNAMP_OK = 0, PNAMP_E_ARG = -1, PNAMP_E_DEVICE = -3;
constexpr int LT = 256;  // threads per workgroup

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

// bfloat16 bits by round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__global__ void k_gather(const float *X, const int *idx, float *out, int B, int d) {
  const int t = blockIdx.x * LT + threadIdx.x;
  if (t >= B * d) return;
  const int b = t / d, j = t - b * d;
  out[t] = X[(size_t)idx[b] * d + j];
}

// ------------------------------------------------------------------ FP32 GEMM on MFMA
constexpr int EPI_FWD = 0, EPI_DATA = 1, EPI_STORE = 2;
struct GemmArgs {
  int M, N, K;                 // C is M x N, K the summed index (this split's range: k0 .. k0 + K)
  const float *A;              // A(m, k) = A[m * sam + k * sak]
  long sam, sak;
  const float *B;              // B(k, n) = B[k * sbk + n * sbn]
  long sbk, sbn;
  float *C;                    // C[m * ldc + n] (+ split * splitc)
  int ldc;
  long splitc;
  int ksplit;                  // K per split (multiple of 16)
  const float *bias;           // EPI_FWD: b[n]
  float *Z;                    // EPI_FWD: pre-activations [m * ldc + n]
  int act;                     // EPI_FWD: SiLU on the output
  const float *Zp;             // EPI_DATA: previous layer's pre-activations [m * ldc + n]
};

// Loads the 64 x 16 panel of A (rows m0.., k0..) and the 16 x 64 panel of B
// into LDS, the contiguous index fastest (coalesced), zero outside the matrix.
__device__ __forceinline__ void stage(const GemmArgs &g, int m0, int n0, int k0, int kend, float (*sa)[17],
                                      float (*sb)[65]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = t + 256 * j;
    int m, k;
    if (g.sak == 1) { k = e & 15; m = e >> 4; }
    else { m = e & 63; k = e >> 6; }
    const int gm = m0 + m, gk = k0 + k;
    sa[m][k] = (gm < g.M && gk < kend) ? g.A[gm * g.sam + gk * g.sak] : 0.0f;
    int n, kb;
    if (g.sbn == 1) { n = e & 63; kb = e >> 6; }
    else { kb = e & 15; n = e >> 4; }
    const int gn = n0 + n, gkb = k0 + kb;
    sb[kb][n] = (gn < g.N && gkb < kend) ? g.B[gkb * g.sbk + gn * g.sbn] : 0.0f;
  }
}

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f32(GemmArgs g) {
  __shared__ float sa[64][17], sb[16][65];  // padded: no bank conflicts on the fragment reads
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int kbeg = blockIdx.z * g.ksplit, kend = min(g.K, kbeg + g.ksplit);
  f32x16 acc;
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    __syncthreads();
    stage(g, m0, n0, k0, kend, sa, sb);
    __syncthreads();
    // lane l: A(wm + l % 32, 2 kk + l / 32), B(2 kk + l / 32, wn + l % 32)
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int k = 2 * kk + (lane >> 5);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sa[wm + (lane & 31)][k], sb[k][wn + (lane & 31)], acc, 0, 0, 0);
    }
  }
  // C(row, col): row = wm + 8 (q / 4) + 4 (lane / 32) + q % 4, col = wn + lane % 32
  const int col = n0 + wn + (lane & 31);
  if (col >= g.N) return;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int row = m0 + wm + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
    if (row >= g.M) continue;
    const size_t o = (size_t)row * g.ldc + col;
    const float c = acc[q];
    if (EPI == EPI_FWD) {
      const float z = c + g.bias[col];
      g.Z[o] = z;
      g.C[o] = g.act ? silu(z) : z;
    } else if (EPI == EPI_DATA) {
      g.C[o] = c * dsilu(g.Zp[o]);
    } else {
      g.C[(size_t)blockIdx.z * g.splitc + o] = c;
    }
  }
}

// dW = sum over the splits of the partials, in split order; db[o] = sum_b d[b][o]
__global__ void k_splitk_sum(const float *P, int nsplit, size_t mn, float *gW, const float *D, int B, int nout,
                             float *gb) {
  const size_t t = (size_t)blockIdx.x * LT + threadIdx.x;
  if (t < mn) {
    float s = 0.0f;
    for (int z = 0; z < nsplit; ++z) s += P[(size_t)z * mn + t];
    gW[t] = s;
  } else if (t < mn + (size_t)nout) {
    const int o = (int)(t - mn);
    float s = 0.0f;
    for (int b = 0; b < B; ++b) s += D[(size_t)b * nout + o];
    gb[o] = s;
  }
}

// ------------------------------------------------------------------ BF16 inference on MFMA
// Weights of layer l in fragment order for the B operand of
// v_mfma_f32_32x32x16_bf16 (B(k, n) = W(n, k)): per 32-column tile nt and
// 16-deep k block kb, 64 lanes x 8 bf16, lane l holding
// W(32 nt + l % 32, 16 kb + 8 (l / 32) + e), e = 0..7; zero padding.
__global__ void k_pack_bf16(const float *W, int nin, int nout, int KB, int NT, unsigned short *out) {
  const int t = blockIdx.x * LT + threadIdx.x;
  const int total = NT * KB * 64 * 8;
  if (t >= total) return;
  const int e = t & 7, lane = (t >> 3) & 63, blk = t >> 9;
  const int kb = blk % KB, nt = blk / KB;
  const int n = 32 * nt + (lane & 31), k = 16 * kb + 8 * (lane >> 5) + e;
  out[t] = (n < nout && k < nin) ? bf16_bits(W[(size_t)n * nin + k]) : (unsigned short)0;
}

struct InferArgs {
  int n, nl;
  int size[PNAMP_MLP_MAX_LAYERS + 1];
  const unsigned short *wpk[PNAMP_MLP_MAX_LAYERS];  // packed bf16 weights per layer
  const float *bias[PNAMP_MLP_MAX_LAYERS];
  const float *X;  // [n][size[0]]
  float *Y;        // [n][size[nl]]
  int ld;          // LDS row stride of the activations (bf16 elements)
};

// 64 samples per workgroup of four waves; a layer's 2 x NT output tiles are
// dealt to the waves, each runs its K blocks on one accumulator and writes
// bias + SiLU back to the other LDS activation buffer as bfloat16.
__global__ __launch_bounds__(256) void k_mlp_infer_bf16(InferArgs a) {
  extern __shared__ unsigned short act[];  // [2][64][ld]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s0 = blockIdx.x * 64;
  unsigned short *cur = act, *nxt = act + 64 * a.ld;
  {
    const int nin = a.size[0], KP = (nin + 15) & ~15;
    for (int e = threadIdx.x; e < 64 * KP; e += 256) {
      const int r = e / KP, k = e - r * KP;
      const int sm = s0 + r;
      cur[r * a.ld + k] = (sm < a.n && k < nin) ? bf16_bits(a.X[(size_t)sm * nin + k]) : (unsigned short)0;
    }
  }
  __syncthreads();
  for (int l = 0; l < a.nl; ++l) {
    const int nin = a.size[l], nout = a.size[l + 1];
    const int KB = (nin + 15) >> 4, NT = (nout + 31) >> 5;
    const bool last = l + 1 == a.nl;
    const bf16x8 *W = (const bf16x8 *)a.wpk[l];
    for (int tile = wave; tile < 2 * NT; tile += 4) {
      const int mt = tile & 1, nt = tile >> 1;
      f32x16 acc;
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
      const unsigned short *arow = cur + (32 * mt + (lane & 31)) * a.ld + 8 * (lane >> 5);
      for (int kb = 0; kb < KB; ++kb) {
        const bf16x8 av = *(const bf16x8 *)(arow + 16 * kb);
        const bf16x8 bv = W[((size_t)nt * KB + kb) * 64 + lane];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
      }
      const int col = 32 * nt + (lane & 31);
      const float b = col < nout ? a.bias[l][col] : 0.0f;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int row = 32 * mt + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        const float z = acc[q] + b;
        if (last) {
          if (col < nout && s0 + row < a.n) a.Y[(size_t)(s0 + row) * nout + col] = z;
        } else {
          nxt[row * a.ld + col] = col < nout ? bf16_bits(silu(z)) : (unsigned short)0;
        }
      }
    }
    __syncthreads();
    unsigned short *t = cur;
    cur = nxt;
    nxt = t;
  }
}

__global__ void k_loss(const float *P, const float *Y, const float *wout, float *D, float *lossp, int B,
                       int nout) {
  __shared__ float red[LT];
  const int t = blockIdx.x * LT + threadIdx.x;
  float l = 0.0f;
  if (t < B * nout) {
    const int o = t % nout;
    const float r = P[t] - Y[t], s = 1.0f / ((float)B * (float)nout);
    D[t] = 2.0f * wout[o] * r * s;
    l = wout[o] * r * r * s;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = LT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, size_t n, float lr, float c1, float c2) {
  const size_t t = (size_t)blockIdx.x * LT + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = 0.9f * m[t] + 0.1f * gt, vt = 0.999f * v[t] + 0.001f * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt * c1) / (sqrtf(vt * c2) + 1e-8f);
}

inline unsigned nblk(size_t n) { return (unsigned)((n + LT - 1) / LT); }

// slices of the batch for the weight gradients: enough workgroups to fill
// the GPU at width 128, partials summed in slice order
inline int ksplits(int B) { return std::max(1, std::min(64, (B + 127) / 128)); }

}  // namespace

struct pnamp_mlp {
  int device = 0;
  int nl = 0;
  int size[PNAMP_MLP_MAX_LAYERS + 1] = {};
  size_t off_w[PNAMP_MLP_MAX_LAYERS] = {}, off_b[PNAMP_MLP_MAX_LAYERS] = {}, npar = 0;
  float *d_par = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
  float *d_X = nullptr, *d_Y = nullptr, *d_wout = nullptr;
  int ndata = 0;
  // per-batch work: activations A[0..nl] (A[0] = inputs), pre-activations Z[1..nl], deltas D[1..nl]
  float *d_A[PNAMP_MLP_MAX_LAYERS + 1] = {}, *d_Z[PNAMP_MLP_MAX_LAYERS + 1] = {}, *d_D[PNAMP_MLP_MAX_LAYERS + 1] = {};
  float *d_Yb = nullptr, *d_lossp = nullptr;
  float *d_part = nullptr;                 // split-K partials of dW [splits][nout][nin]
  unsigned short *d_wpk = nullptr;         // bf16 weights in fragment order (inference)
  size_t off_pk[PNAMP_MLP_MAX_LAYERS] = {};
  bool pk_valid = false;
  int *d_idx = nullptr;
  int bcap = 0;
  long step = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  float last_ms = 0.0f;
  std::string err;
};

#define MCHK(m, x)                                                          \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      (m)->err = std::string(#x) + ": " + hipGetErrorString(e_);            \
      return PNAMP_E_DEVICE;                                                \
    }                                                                       \
  } while (0)

static int mfail(pnamp_mlp *m, int code, const std::string &msg) {
  m->err = msg;
  return code;
}

static int mrealloc(pnamp_mlp *m, void **p, size_t bytes) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  MCHK(m, hipMalloc(p, std::max<size_t>(bytes, 4)));
  return 0;
}

// Work buffers for batches of up to B samples.
static int ensure_batch(pnamp_mlp *m, int B) {
  if (B <= m->bcap) return 0;
  int r = 0;
  for (int l = 0; l <= m->nl; ++l) {
    const size_t bytes = sizeof(float) * (size_t)B * m->size[l];
    if ((r = mrealloc(m, (void **)&m->d_A[l], bytes)) || (r = mrealloc(m, (void **)&m->d_Z[l], bytes)) ||
        (r = mrealloc(m, (void **)&m->d_D[l], bytes)))
      return r;
  }
  const size_t nout = m->size[m->nl];
  size_t wmax = 0;
  for (int l = 0; l < m->nl; ++l) wmax = std::max(wmax, (size_t)m->size[l] * m->size[l + 1]);
  if ((r = mrealloc(m, (void **)&m->d_Yb, sizeof(float) * B * nout)) ||
      (r = mrealloc(m, (void **)&m->d_lossp, sizeof(float) * nblk((size_t)B * nout))) ||
      (r = mrealloc(m, (void **)&m->d_idx, sizeof(int) * B)) ||
      (r = mrealloc(m, (void **)&m->d_part, sizeof(float) * wmax * ksplits(B))))
    return r;
  m->bcap = B;
  return 0;
}

static GemmArgs gemm_args() {
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  return g;
}

static dim3 gemm_grid(const GemmArgs &g, int splits) {
  return dim3((unsigned)((g.N + 63) / 64), (unsigned)((g.M + 63) / 64), (unsigned)splits);
}

// Forward pass over the batch in A[0] (FP32 on MFMA); Z kept for backpropagation.
static void forward(pnamp_mlp *m, int B) {
  for (int l = 0; l < m->nl; ++l) {
    const int nin = m->size[l], nout = m->size[l + 1];
    GemmArgs g = gemm_args();
    g.M = B; g.N = nout; g.K = nin; g.ksplit = (nin + 15) & ~15;
    g.A = m->d_A[l]; g.sam = nin; g.sak = 1;                    // a(b, i)
    g.B = m->d_par + m->off_w[l]; g.sbk = 1; g.sbn = nin;       // W^T(i, o) = W[o][i]
    g.C = m->d_A[l + 1]; g.ldc = nout;
    g.bias = m->d_par + m->off_b[l]; g.Z = m->d_Z[l + 1]; g.act = (l + 1 < m->nl) ? 1 : 0;
    hipLaunchKernelGGL(k_gemm_f32<EPI_FWD>, gemm_grid(g, 1), dim3(256), 0, m->stream, g);
  }
}

// Loss block sums into d_lossp and the gradient into d_grad for the batch in
// A[0] with targets d_Yb.
static void loss_backward(pnamp_mlp *m, int B) {
  const int L = m->nl, nout = m->size[L];
  hipLaunchKernelGGL(k_loss, dim3(nblk((size_t)B * nout)), dim3(LT), 0, m->stream, m->d_A[L], m->d_Yb, m->d_wout,
                     m->d_D[L], m->d_lossp, B, nout);
  const int S = ksplits(B), kc = ((B + S - 1) / S + 15) & ~15;
  for (int l = L - 1; l >= 0; --l) {
    const int nin = m->size[l], no = m->size[l + 1];
    // dW = d^T a_prev over the batch, in S fixed slices of the batch, then summed
    GemmArgs g = gemm_args();
    g.M = no; g.N = nin; g.K = B; g.ksplit = kc;
    g.A = m->d_D[l + 1]; g.sam = 1; g.sak = no;                 // d^T(o, b)
    g.B = m->d_A[l]; g.sbk = nin; g.sbn = 1;                    // a_prev(b, i)
    g.C = m->d_part; g.ldc = nin; g.splitc = (long)no * nin;
    hipLaunchKernelGGL(k_gemm_f32<EPI_STORE>, gemm_grid(g, (B + kc - 1) / kc), dim3(256), 0, m->stream, g);
    const size_t mn = (size_t)no * nin;
    hipLaunchKernelGGL(k_splitk_sum, dim3(nblk(mn + no)), dim3(LT), 0, m->stream, m->d_part, (B + kc - 1) / kc, mn,
                       m->d_grad + m->off_w[l], m->d_D[l + 1], B, no, m->d_grad + m->off_b[l]);
    if (l > 0) {  // d_prev = (d W) * silu'(z_prev)
      GemmArgs h = gemm_args();
      h.M = B; h.N = nin; h.K = no; h.ksplit = (no + 15) & ~15;
      h.A = m->d_D[l + 1]; h.sam = no; h.sak = 1;               // d(b, o)
      h.B = m->d_par + m->off_w[l]; h.sbk = nin; h.sbn = 1;     // W(o, i)
      h.C = m->d_D[l]; h.ldc = nin; h.Zp = m->d_Z[l];
      hipLaunchKernelGGL(k_gemm_f32<EPI_DATA>, gemm_grid(h, 1), dim3(256), 0, m->stream, h);
    }
  }
}

extern "C" int pnamp_mlp_create(int nlayers, const int32_t *sizes, uint32_t seed, int device, pnamp_mlp **out) {
  if (!out) return PNAMP_E_ARG;
  *out = nullptr;
  if (!sizes || nlayers < 1 || nlayers > PNAMP_MLP_MAX_LAYERS) return PNAMP_E_ARG;
  for (int l = 0; l <= nlayers; ++l)
    if (sizes[l] < 1 || sizes[l] > PNAMP_MLP_MAX_WIDTH) return PNAMP_E_ARG;
  pnamp_mlp *m = new pnamp_mlp();
  *out = m;
  m->device = device;
  m->nl = nlayers;
  size_t off = 0;
  for (int l = 0; l <= nlayers; ++l) m->size[l] = sizes[l];
  for (int l = 0; l < nlayers; ++l) {
    m->off_w[l] = off;
    off += (size_t)sizes[l + 1] * sizes[l];
    m->off_b[l] = off;
    off += sizes[l + 1];
  }
  m->npar = off;
  MCHK(m, hipSetDevice(device));
  MCHK(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  MCHK(m, hipEventCreate(&m->ev[0]));
  MCHK(m, hipEventCreate(&m->ev[1]));
  std::vector<float> p(off, 0.0f);
  std::mt19937 rng(seed);
  for (int l = 0; l < nlayers; ++l) {
    const float a = std::sqrt(6.0f / (float)(sizes[l] + sizes[l + 1]));
    std::uniform_real_distribution<float> u(-a, a);
    for (size_t i = 0; i < (size_t)sizes[l + 1] * sizes[l]; ++i) p[m->off_w[l] + i] = u(rng);
  }
  int r = 0;
  if ((r = mrealloc(m, (void **)&m->d_par, off * 4)) || (r = mrealloc(m, (void **)&m->d_grad, off * 4)) ||
      (r = mrealloc(m, (void **)&m->d_m, off * 4)) || (r = mrealloc(m, (void **)&m->d_v, off * 4)))
    return r;
  MCHK(m, hipMemcpy(m->d_par, p.data(), off * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemset(m->d_m, 0, off * 4));
  MCHK(m, hipMemset(m->d_v, 0, off * 4));
  return PNAMP_OK;
}

extern "C" int64_t pnamp_mlp_n_params(pnamp_mlp *m) { return m ? (int64_t)m->npar : PNAMP_E_ARG; }

extern "C" int pnamp_mlp_get_params(pnamp_mlp *m, float *params) {
  if (!m || !params) return PNAMP_E_ARG;
  MCHK(m, hipMemcpy(params, m->d_par, m->npar * 4, hipMemcpyDeviceToHost));
  return PNAMP_OK;
}

extern "C" int pnamp_mlp_set_params(pnamp_mlp *m, const float *params) {
  if (!m || !params) return PNAMP_E_ARG;
  MCHK(m, hipMemcpy(m->d_par, params, m->npar * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemset(m->d_m, 0, m->npar * 4));
  MCHK(m, hipMemset(m->d_v, 0, m->npar * 4));
  m->step = 0;
  m->pk_valid = false;
  return PNAMP_OK;
}

extern "C" int pnamp_mlp_set_data(pnamp_mlp *m, int n, const float *X, const float *Y, const float *wout) {
  if (!m || n < 1 || !X || !Y) return PNAMP_E_ARG;
  const size_t nin = m->size[0], nout = m->size[m->nl];
  int r = 0;
  if ((r = mrealloc(m, (void **)&m->d_X, 4 * n * nin)) || (r = mrealloc(m, (void **)&m->d_Y, 4 * n * nout)) ||
      (r = mrealloc(m, (void **)&m->d_wout, 4 * nout)))
    return r;
  MCHK(m, hipMemcpy(m->d_X, X, 4 * n * nin, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->d_Y, Y, 4 * n * nout, hipMemcpyHostToDevice));
  std::vector<float> w(nout, 1.0f);
  if (wout) std::copy(wout, wout + nout, w.begin());
  MCHK(m, hipMemcpy(m->d_wout, w.data(), 4 * nout, hipMemcpyHostToDevice));
  m->ndata = n;
  return PNAMP_OK;
}

// Gathers samples idx into the batch buffers (A[0], Yb).
static int load_batch(pnamp_mlp *m, int B) {
  const int nin = m->size[0], nout = m->size[m->nl];
  hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)B * nin)), dim3(LT), 0, m->stream, m->d_X, m->d_idx, m->d_A[0], B,
                     nin);
  hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)B * nout)), dim3(LT), 0, m->stream, m->d_Y, m->d_idx, m->d_Yb, B,
                     nout);
  MCHK(m, hipGetLastError());
  return 0;
}

extern "C" int pnamp_mlp_loss_grad(pnamp_mlp *m, int n, const int32_t *idx, float *loss, float *grad) {
  if (!m || n < 1 || !loss) return PNAMP_E_ARG;
  if (n > m->ndata) return mfail(m, PNAMP_E_ARG, "more samples than the training set holds");
  std::vector<int> ix(n);
  for (int i = 0; i < n; ++i) {
    ix[i] = idx ? idx[i] : i;
    if (ix[i] < 0 || ix[i] >= m->ndata) return mfail(m, PNAMP_E_ARG, "sample index out of range");
  }
  int r = ensure_batch(m, n);
  if (r) return r;
  MCHK(m, hipMemcpyAsync(m->d_idx, ix.data(), 4 * n, hipMemcpyHostToDevice, m->stream));
  if ((r = load_batch(m, n))) return r;
  forward(m, n);
  loss_backward(m, n);
  MCHK(m, hipGetLastError());
  const size_t nb = nblk((size_t)n * m->size[m->nl]);
  std::vector<float> lp(nb);
  MCHK(m, hipMemcpyAsync(lp.data(), m->d_lossp, 4 * nb, hipMemcpyDeviceToHost, m->stream));
  if (grad) MCHK(m, hipMemcpyAsync(grad, m->d_grad, 4 * m->npar, hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  double s = 0.0;
  for (float x : lp) s += x;
  *loss = (float)s;
  return PNAMP_OK;
}

extern "C" int pnamp_mlp_train(pnamp_mlp *m, int epochs, int batch, float lr, float decay, uint32_t seed,
                               float *epoch_loss) {
  if (!m || epochs < 1 || batch < 1) return PNAMP_E_ARG;
  if (m->ndata < 1) return mfail(m, PNAMP_E_ARG, "no training data");
  const int B = std::min(batch, m->ndata);
  int r = ensure_batch(m, B);
  if (r) return r;
  std::mt19937 rng(seed);
  std::vector<int> perm(m->ndata);
  for (int i = 0; i < m->ndata; ++i) perm[i] = i;
  const int nbat = m->ndata / B;  // whole batches only; the tail joins the next epoch's shuffle
  const size_t nb = nblk((size_t)B * m->size[m->nl]);
  std::vector<float> lp(nb);
  int *d_perm = nullptr;
  MCHK(m, hipMalloc((void **)&d_perm, 4 * (size_t)m->ndata));
  struct Free {
    int *p;
    ~Free() { if (p) (void)hipFree(p); }
  } fr{d_perm};
  float lr_e = lr;
  for (int e = 0; e < epochs; ++e, lr_e *= decay) {
    std::shuffle(perm.begin(), perm.end(), rng);
    MCHK(m, hipMemcpyAsync(d_perm, perm.data(), 4 * (size_t)m->ndata, hipMemcpyHostToDevice, m->stream));
    double esum = 0.0;
    for (int bi = 0; bi < nbat; ++bi) {
      MCHK(m, hipMemcpyAsync(m->d_idx, d_perm + (size_t)bi * B, 4 * (size_t)B, hipMemcpyDeviceToDevice, m->stream));
      if ((r = load_batch(m, B))) return r;
      forward(m, B);
      loss_backward(m, B);
      ++m->step;
      m->pk_valid = false;
      const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m->step));
      const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m->step));
      hipLaunchKernelGGL(k_adam, dim3(nblk(m->npar)), dim3(LT), 0, m->stream, m->d_par, m->d_grad, m->d_m, m->d_v,
                         m->npar, lr_e, c1, c2);
      MCHK(m, hipGetLastError());
      if (epoch_loss) {
        MCHK(m, hipMemcpyAsync(lp.data(), m->d_lossp, 4 * nb, hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
        for (float x : lp) esum += x;
      }
    }
    if (epoch_loss) epoch_loss[e] = (float)(esum / std::max(nbat, 1));
  }
  MCHK(m, hipStreamSynchronize(m->stream));
  return PNAMP_OK;
}

// bf16 copies of the weights in fragment order (after any change of the parameters)
static int pack_bf16(pnamp_mlp *m) {
  if (m->pk_valid) return 0;
  size_t tot = 0;
  for (int l = 0; l < m->nl; ++l) {
    m->off_pk[l] = tot;
    tot += (size_t)((m->size[l + 1] + 31) / 32) * ((m->size[l] + 15) / 16) * 512;
  }
  if (!m->d_wpk) {
    int r = mrealloc(m, (void **)&m->d_wpk, 2 * tot);
    if (r) return r;
  }
  for (int l = 0; l < m->nl; ++l) {
    const int KB = (m->size[l] + 15) / 16, NT = (m->size[l + 1] + 31) / 32;
    hipLaunchKernelGGL(k_pack_bf16, dim3(nblk((size_t)NT * KB * 512)), dim3(LT), 0, m->stream,
                       m->d_par + m->off_w[l], m->size[l], m->size[l + 1], KB, NT, m->d_wpk + m->off_pk[l]);
  }
  MCHK(m, hipGetLastError());
  m->pk_valid = true;
  return 0;
}

extern "C" int pnamp_mlp_predict(pnamp_mlp *m, int n, const float *X, float *Y, int bf16) {
  if (!m || n < 1 || !X || !Y) return PNAMP_E_ARG;
  const int nin = m->size[0], nout = m->size[m->nl];
  const int B = std::min(n, 1 << 16);
  int r = ensure_batch(m, B);
  if (r) return r;
  if (bf16 && (r = pack_bf16(m))) return r;
  InferArgs ia;
  std::memset(&ia, 0, sizeof(ia));
  int wmax = 16;
  if (bf16) {
    ia.nl = m->nl;
    for (int l = 0; l <= m->nl; ++l) {
      ia.size[l] = m->size[l];
      wmax = std::max(wmax, (m->size[l] + 31) & ~31);
    }
    for (int l = 0; l < m->nl; ++l) {
      ia.wpk[l] = m->d_wpk + m->off_pk[l];
      ia.bias[l] = m->d_par + m->off_b[l];
    }
    ia.ld = wmax + 8;  // padded rows: the fragment reads of 32 rows fall in different banks
  }
  const size_t lds = sizeof(unsigned short) * 2 * 64 * (size_t)(wmax + 8);
  float ms_total = 0.0f;
  for (int s = 0; s < n; s += B) {
    const int b = std::min(B, n - s);
    MCHK(m, hipMemcpyAsync(m->d_A[0], X + (size_t)s * nin, 4 * (size_t)b * nin, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipEventRecord(m->ev[0], m->stream));
    if (bf16) {
      ia.n = b;
      ia.X = m->d_A[0];
      ia.Y = m->d_A[m->nl];
      hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((b + 63) / 64), dim3(256), lds, m->stream, ia);
    } else {
      forward(m, b);
    }
    MCHK(m, hipEventRecord(m->ev[1], m->stream));
    MCHK(m, hipGetLastError());
    MCHK(m, hipMemcpyAsync(Y + (size_t)s * nout, m->d_A[m->nl], 4 * (size_t)b * nout, hipMemcpyDeviceToHost,
                           m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    float ms = 0.0f;
    MCHK(m, hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
    ms_total += ms;
  }
  m->last_ms = ms_total;
  return PNAMP_OK;
}

extern "C" float pnamp_mlp_last_ms(pnamp_mlp *m) { return m ? m->last_ms : -1.0f; }

extern "C" const char *pnamp_mlp_last_error(pnamp_mlp *m) { return m ? m->err.c_str() : "null network"; }

extern "C" void pnamp_mlp_destroy(pnamp_mlp *m) {
  if (!m) return;
  float *bufs[] = {m->d_par, m->d_grad, m->d_m, m->d_v, m->d_X, m->d_Y, m->d_wout, m->d_Yb, m->d_lossp, m->d_part};
  for (float *b : bufs)
    if (b) (void)hipFree(b);
  for (int l = 0; l <= PNAMP_MLP_MAX_LAYERS; ++l) {
    if (m->d_A[l]) (void)hipFree(m->d_A[l]);
    if (m->d_Z[l]) (void)hipFree(m->d_Z[l]);
    if (m->d_D[l]) (void)hipFree(m->d_D[l]);
  }
  if (m->d_idx) (void)hipFree(m->d_idx);
  if (m->d_wpk) (void)hipFree(m->d_wpk);
  for (auto e : m->ev)
    if (e) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}
