// Learned surrogate: a fully connected network (SiLU hidden layers, linear
// output, FP32 parameters) with hand-written HIP kernels for the forward pass,
// the loss, backpropagation and Adam, plus a bfloat16 inference path.  C-ABI
// in include/pnamp_mlp.h.
//
// Per layer l with inputs a (B x n_in) and outputs z = W a + b (B x n_out):
//   k_dense_fwd      z and a' = silu(z) (identity on the last layer), one
//                    thread per (sample, output)
//   k_loss           d = 2 w_o (f - y) / (B n_out) and block sums of the loss
//   k_dense_bwd_data d_prev = (W^T d) * silu'(z_prev)
//   k_dense_bwd_w    dW = d^T a_prev, db = sum_b d: a workgroup per output
//                    row, a thread per input column, the batch in a loop
//   k_adam           the Adam update over the flat parameter vector
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

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

__device__ __forceinline__ float to_bf16(float f) {  // round to nearest even, back to FP32
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xFFFF0000u);
}

__global__ void k_gather(const float *X, const int *idx, float *out, int B, int d) {
  const int t = blockIdx.x * LT + threadIdx.x;
  if (t >= B * d) return;
  const int b = t / d, j = t - b * d;
  out[t] = X[(size_t)idx[b] * d + j];
}

__global__ void k_dense_fwd(const float *W, const float *bias, const float *A, float *Z, float *Aout, int B,
                            int nin, int nout, int act, int bf16) {
  const int t = blockIdx.x * LT + threadIdx.x;
  if (t >= B * nout) return;
  const int b = t / nout, o = t - b * nout;
  const float *w = W + (size_t)o * nin, *a = A + (size_t)b * nin;
  float z = bias[o];
  if (bf16) {
    for (int i = 0; i < nin; ++i) z = fmaf(to_bf16(w[i]), to_bf16(a[i]), z);
  } else {
    for (int i = 0; i < nin; ++i) z = fmaf(w[i], a[i], z);
  }
  if (Z) Z[t] = z;
  Aout[t] = act ? z * sigm(z) : z;
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

__global__ void k_dense_bwd_data(const float *W, const float *D, const float *Zp, float *Dp, int B, int nin,
                                 int nout) {
  const int t = blockIdx.x * LT + threadIdx.x;
  if (t >= B * nin) return;
  const int b = t / nin, i = t - b * nin;
  const float *d = D + (size_t)b * nout;
  float g = 0.0f;
  for (int o = 0; o < nout; ++o) g = fmaf(W[(size_t)o * nin + i], d[o], g);
  const float z = Zp[t], s = sigm(z);
  Dp[t] = g * s * (1.0f + z * (1.0f - s));
}

__global__ void k_dense_bwd_w(const float *D, const float *Ap, float *gW, float *gb, int B, int nin, int nout) {
  const int o = blockIdx.x, i = threadIdx.x;
  if (i > nin) return;
  float g = 0.0f;
  if (i < nin) {
    for (int b = 0; b < B; ++b) g = fmaf(D[(size_t)b * nout + o], Ap[(size_t)b * nin + i], g);
    gW[(size_t)o * nin + i] = g;
  } else {
    for (int b = 0; b < B; ++b) g += D[(size_t)b * nout + o];
    gb[o] = g;
  }
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
  if ((r = mrealloc(m, (void **)&m->d_Yb, sizeof(float) * B * nout)) ||
      (r = mrealloc(m, (void **)&m->d_lossp, sizeof(float) * nblk((size_t)B * nout))) ||
      (r = mrealloc(m, (void **)&m->d_idx, sizeof(int) * B)))
    return r;
  m->bcap = B;
  return 0;
}

// Forward pass over the batch in A[0]; Z kept for backpropagation.
static void forward(pnamp_mlp *m, int B, int bf16) {
  for (int l = 0; l < m->nl; ++l) {
    const int nin = m->size[l], nout = m->size[l + 1];
    const int act = (l + 1 < m->nl) ? 1 : 0;
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((size_t)B * nout)), dim3(LT), 0, m->stream, m->d_par + m->off_w[l],
                       m->d_par + m->off_b[l], m->d_A[l], m->d_Z[l + 1], m->d_A[l + 1], B, nin, nout, act, bf16);
  }
}

// Loss block sums into d_lossp and the gradient into d_grad for the batch in
// A[0] with targets d_Yb.
static void loss_backward(pnamp_mlp *m, int B) {
  const int L = m->nl, nout = m->size[L];
  hipLaunchKernelGGL(k_loss, dim3(nblk((size_t)B * nout)), dim3(LT), 0, m->stream, m->d_A[L], m->d_Yb, m->d_wout,
                     m->d_D[L], m->d_lossp, B, nout);
  for (int l = L - 1; l >= 0; --l) {
    const int nin = m->size[l], no = m->size[l + 1];
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(no), dim3((nin + 1 + 63) / 64 * 64), 0, m->stream, m->d_D[l + 1],
                       m->d_A[l], m->d_grad + m->off_w[l], m->d_grad + m->off_b[l], B, nin, no);
    if (l > 0)
      hipLaunchKernelGGL(k_dense_bwd_data, dim3(nblk((size_t)B * nin)), dim3(LT), 0, m->stream,
                         m->d_par + m->off_w[l], m->d_D[l + 1], m->d_Z[l], m->d_D[l], B, nin, no);
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
  forward(m, n, 0);
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
      forward(m, B, 0);
      loss_backward(m, B);
      ++m->step;
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

extern "C" int pnamp_mlp_predict(pnamp_mlp *m, int n, const float *X, float *Y, int bf16) {
  if (!m || n < 1 || !X || !Y) return PNAMP_E_ARG;
  const int nin = m->size[0], nout = m->size[m->nl];
  const int B = std::min(n, 1 << 16);
  int r = ensure_batch(m, B);
  if (r) return r;
  float ms_total = 0.0f;
  for (int s = 0; s < n; s += B) {
    const int b = std::min(B, n - s);
    MCHK(m, hipMemcpyAsync(m->d_A[0], X + (size_t)s * nin, 4 * (size_t)b * nin, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipEventRecord(m->ev[0], m->stream));
    forward(m, b, bf16);
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
  float *bufs[] = {m->d_par, m->d_grad, m->d_m, m->d_v, m->d_X, m->d_Y, m->d_wout, m->d_Yb, m->d_lossp};
  for (float *b : bufs)
    if (b) (void)hipFree(b);
  for (int l = 0; l <= PNAMP_MLP_MAX_LAYERS; ++l) {
    if (m->d_A[l]) (void)hipFree(m->d_A[l]);
    if (m->d_Z[l]) (void)hipFree(m->d_Z[l]);
    if (m->d_D[l]) (void)hipFree(m->d_D[l]);
  }
  if (m->d_idx) (void)hipFree(m->d_idx);
  for (auto e : m->ev)
    if (e) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}
