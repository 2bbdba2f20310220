// pnamp_gpu.hip -- MI355X (gfx950) implementation of the PNAMP projected-kernel
// path: FP64 rotation, Pfaffian norm overlap, transition densities, energy
// kernel, gauge (particle-number) sum, J-projection and Hill-Wheeler solve,
// behind the C-ABI of include/pnamp_gpu.h.
//
// Work decomposition (one GPU):
//   k_dtables     Wigner d tables: rotation blocks d^j_{m m'}(beta) and
//                 projection d^J_{K K'}(beta), d^J_{K K'}(pi - beta), per beta.
//   k_state_prep  per HFB state: U^{-1} (Gauss-Jordan), Z = V U^{-1},
//                 vacuum overlap det(1 + Z^T Z)^{-1/4} (Cholesky).
//   k_plain_*     plain form (plain_gpu.hip): per (pair, Euler point) the
//                 gauge-independent rotation products, then one workgroup
//                 per gauge point: Pfaffian of the 2n x 2n skew matrix (full
//                 pivoting, the ZPfaffianF algorithm), T^{-1} applied by
//                 Gauss-Jordan, rho / kappa, Tr rho per species, E_kin, <J^2>
//                 and the two-body energy from the regrouped m-scheme list;
//                 then the gauge sum into ZRES(pair, Euler, obs).
//   k_fac_*       factored form for states without proton-neutron mixing.
//   k_project     one workgroup per (pair, observable): J-projection sums
//                 over the Euler mesh (plain or literal fold).
//   k_hw_*        Hill-Wheeler: Jacobi eigen-solves of N and N^-1/2 H N^-1/2.
//
// Everything is FP64.  No CPU fallback: every numerical step of a kernel
// call runs on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pnamp_gpu.h"
#include "plain_gpu.h"
#include "pnamp_dev.h"

namespace {

using namespace pnamp_dev;

constexpr int NMAX = PNAMP_NMAX;      // largest n (m-states, both species)
constexpr int HMAX = NMAX / 2;        // largest species half
constexpr int FAC_HMAX = 22;          // largest species half of the factored and mixed forms
constexpr int RPACK_MAX = 256;        // packed rotation-block entries per species
constexpr int NOB2_MAX = 2 * 8 * 8;  // orbit pairs of both species (<= 8 orbits per species)
constexpr int JMAX_MAX = 12;          // largest projected J supported
constexpr int NOBS = PNAMP_NOBS;

__constant__ double c_fact[40];

// d^j_{m1 m2}(beta) = <j m1|exp(-i beta J_y)|j m2>, doubled arguments.
__device__ double wigner_d_dev(int tj, int tm1, int tm2, double beta) {
  if (abs(tm1) > tj || abs(tm2) > tj || ((tj + tm1) & 1) || ((tj + tm2) & 1)) return 0.0;
  const int jpm1 = (tj + tm1) / 2, jmm1 = (tj - tm1) / 2, jpm2 = (tj + tm2) / 2, jmm2 = (tj - tm2) / 2;
  const double pre = sqrt(c_fact[jpm1] * c_fact[jmm1] * c_fact[jpm2] * c_fact[jmm2]);
  double s, c;
  sincos(0.5 * beta, &s, &c);
  const int dm = (tm1 - tm2) / 2;
  const int smin = dm < 0 ? -dm : 0;
  const int smax = jpm2 < jmm1 ? jpm2 : jmm1;
  double sum = 0.0;
  for (int k = smin; k <= smax; ++k) {
    const double den = c_fact[jpm2 - k] * c_fact[k] * c_fact[dm + k] * c_fact[jmm1 - k];
    double term = 1.0;
    for (int e = 0; e < tj - dm - 2 * k; ++e) term *= c;
    for (int e = 0; e < dm + 2 * k; ++e) term *= s;
    term /= den;
    sum += ((dm + k) & 1) ? -term : term;
  }
  return pre * sum;
}

// ------------------------------------------------------------------ tables
struct TableParams {
  int nbet, h, npack, jmax, npj;
  const double *bet;
  const int *pack_a, *pack_b;  // local (species-half) indices of packed entries
  const int *twoj, *twom;      // species-half basis
  double *rtab;                // [nbet][npack]
  double *ptab;                // [nbet][2][npj]
};

__global__ void k_dtables(TableParams p) {
  const int ib = blockIdx.x;
  const double bet = p.bet[ib];
  for (int e = threadIdx.x; e < p.npack; e += blockDim.x) {
    const int a = p.pack_a[e], b = p.pack_b[e];
    p.rtab[(size_t)ib * p.npack + e] = wigner_d_dev(p.twoj[a], p.twom[a], p.twom[b], bet);
  }
  for (int e = threadIdx.x; e < p.npj; e += blockDim.x) {
    int J = 0, off = 0;
    while (off + (2 * J + 1) * (2 * J + 1) <= e) {
      off += (2 * J + 1) * (2 * J + 1);
      ++J;
    }
    const int r = e - off, K1 = r / (2 * J + 1) - J, K2 = r % (2 * J + 1) - J;
    p.ptab[((size_t)ib * 2 + 0) * p.npj + e] = wigner_d_dev(2 * J, 2 * K1, 2 * K2, bet);
    p.ptab[((size_t)ib * 2 + 1) * p.npj + e] = wigner_d_dev(2 * J, 2 * K1, 2 * K2, M_PI - bet);
  }
}

// ------------------------------------------------------------------ state prep
// One workgroup per state.  zbareover restated for real U, V:
// Z = V U^{-1}; det(1 + Z^T Z)^{-1/4} equals exp(-sum log(1 + w)/4) over the
// eigenvalues w of Z^T Z (real symmetric, non-negative).
struct PrepParams {
  int n;
  const double *U, *V;
  double *Z, *det;
  int *status;  // [nstates]
};

__global__ __launch_bounds__(256) void k_state_prep(PrepParams p) {
  extern __shared__ double sprep[];  // A, B (n x n each), fcol (n)
  __shared__ int piv;
  const int q = blockIdx.x, n = p.n, nn = n * n, tid = threadIdx.x;
  double *const A = sprep, *const B = sprep + nn, *const fcol = sprep + 2 * nn;
  const double *U = p.U + (size_t)q * nn, *V = p.V + (size_t)q * nn;
  double *Z = p.Z + (size_t)q * nn;
  int status = 0;
  for (int e = tid; e < nn; e += blockDim.x) {
    A[e] = U[e];
    B[e] = ((e % n) == (e / n)) ? 1.0 : 0.0;
  }
  __syncthreads();
  // Gauss-Jordan with partial pivoting: A -> 1, B -> U^{-1}
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      int pr = k;
      double m = fabs(A[k + k * n]);
      for (int i = k + 1; i < n; ++i)
        if (fabs(A[i + k * n]) > m) { m = fabs(A[i + k * n]); pr = i; }
      piv = pr;
    }
    __syncthreads();
    const int pr = piv;
    if (pr != k)
      for (int j = tid; j < 2 * n; j += blockDim.x) {
        double *M = j < n ? A : B;
        const int jj = j < n ? j : j - n;
        const double t = M[k + jj * n];
        M[k + jj * n] = M[pr + jj * n];
        M[pr + jj * n] = t;
      }
    __syncthreads();
    const double pv = A[k + k * n];
    if (pv == 0.0) status = k + 1;
    const double ip = 1.0 / pv;
    for (int i = tid; i < n; i += blockDim.x) fcol[i] = (i == k) ? 0.0 : A[i + k * n];
    for (int j = tid; j < 2 * n; j += blockDim.x) {
      double *M = j < n ? A : B;
      const int jj = j < n ? j : j - n;
      M[k + jj * n] *= ip;
    }
    __syncthreads();
    for (int e = tid; e < 2 * nn; e += blockDim.x) {
      const int i = e % n, jj = (e / n) % n;
      double *M = e < nn ? A : B;
      if (i != k) M[i + jj * n] -= fcol[i] * M[k + jj * n];
    }
    __syncthreads();
  }
  // Z = V U^{-1} -> global and A
  for (int e = tid; e < nn; e += blockDim.x) {
    const int i = e % n, j = e / n;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s = fma(V[i + k * n], B[k + j * n], s);
    Z[e] = s;
    A[e] = s;
  }
  __syncthreads();
  // B = 1 + Z^T Z
  for (int e = tid; e < nn; e += blockDim.x) {
    const int i = e % n, j = e / n;
    double s = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k) s = fma(A[k + i * n], A[k + j * n], s);
    B[e] = s;
  }
  __syncthreads();
  // Cholesky (lower, in B), log det = 2 sum log L_kk
  double logdet = 0.0;
  for (int k = 0; k < n; ++k) {
    const double d = sqrt(B[k + k * n]);
    logdet += 2.0 * log(d);
    __syncthreads();
    for (int i = k + 1 + tid; i < n; i += blockDim.x) B[i + k * n] /= d;
    __syncthreads();
    for (int e = tid; e < nn; e += blockDim.x) {
      const int i = e % n, j = e / n;
      if (j > k && i >= j) B[i + j * n] -= B[i + k * n] * B[j + k * n];
    }
    __syncthreads();
  }
  if (tid == 0) {
    p.det[q] = exp(-logdet / 4.0);
    p.status[q] = status;
  }
}

// One m-scheme list entry on the host: V, the four 0-based indices packed
// a | b << 8 | c << 16 | d << 24, and a flag word.
struct VEntry {
  double me;
  uint32_t idx;
  uint32_t pad;
};

// ------------------------------------------------------------------ factored form
// For states without proton-neutron mixing (U, V block-diagonal by species
// after ordering the quasiparticles) every matrix of the path is
// block-diagonal: the 2n x 2n Pfaffian is (-1)^h Pf(S_p) Pf(S_n) with
// S_s = [[e^{2 i phi_s} R Z1s R^T, -1], [1, -Z0s]] (2h x 2h), and the species-s
// densities depend only on phi_s, through T_s = e^{i phi_s} X_s + e^{-i phi_s} Y_s
// with X_s = V0s^T R V1s, Y_s = U0s^T R* U1s fixed per Euler point.  So each
// Euler point needs nphi Pfaffians / inverses / density sets per species
// instead of nphi^2 full ones.  The energy splits into E_pp(phi_p) + E_nn(phi_n)
// (species lists) + E_pn(phi_p, phi_n), a bilinear form in rho_p and rho_n
// evaluated for all nphi^2 gauge pairs by k_fac_combine.
//
// Four launches per chunk of pairs:
//   k_fac_setup    one workgroup per (pair, Euler point, species): the
//                  gauge-independent X, Y, R V1s, R* U1s, M0 = R Z1s R^T -> HBM;
//   k_fac_phi<h>   one wavefront per (pair, Euler point, species, phi), twice
//                  over the same grid: the Pfaffian (the skew matrix in a 2D
//                  block-cyclic register layout), then Gauss-Jordan (one
//                  column per lane) and the densities, rho / kappa in LDS for
//                  the list; no workgroup barriers;
//   k_fac_combine  one workgroup per (pair, Euler point): the p-n bilinear form
//                  and the nphi^2 gauge sum.
constexpr int NTS = 256;     // threads of the setup workgroup
constexpr int NTC = 512;     // threads of the combine workgroup: one G row per thread (h^2 <= 484),
                              // so the longest row (pf: 60 entries) bounds the pass; eight waves
constexpr int NPHI_MAX = 16;
constexpr int FAC_OUT = 9;   // per (pair, Euler, species, phi): Pf, E2, Ekin, Tr rho, A, t_+, t_-, t_z, E2 (kappa part, mixed forms)
constexpr int MIX_FP64 = 0, MIX_FP32 = 1, MIX_BF16 = 2;  // precision of the V contraction

struct SetupParams {
  int n, h, nalp, ngam, npack, NE;
  const double *alp, *gam;
  const double *Uf, *Vf;            // [nstates][2][h*h], species blocks, column-major
  const double *Z;                  // [nstates][n*n]
  const double *rtab;               // [nbet][npack]
  const int *pk_a, *pk_b, *blk_off, *blk_sz, *blk_base, *twom;
  const int *q0, *q1;
  double2 *sbuf;                    // [npairs][NE][2][5][h*h]: X, Y, R V1s, R* U1s, M0 (column-major)
};

__global__ __launch_bounds__(NTS) void k_fac_setup(SetupParams p) {
  extern __shared__ double2 sm[];  // 3 h^2: R V1s, R* U1s, R Z1s
  __shared__ double2 sR[RPACK_MAX];
  __shared__ short m_off[HMAX], m_sz[HMAX], m_base[HMAX];
  const int tid = threadIdx.x;
  const int iang = blockIdx.x, ip = blockIdx.y, sp = blockIdx.z;
  const int n = p.n, h = p.h, hh = h * h;
  const int ig = iang % p.ngam, ia = (iang / p.ngam) % p.nalp, ib = iang / (p.ngam * p.nalp);
  const double alp = p.alp[ia], gam = p.gam[ig];
  const int q0 = p.q0[ip], q1 = p.q1[ip];
  const double *U0 = p.Uf + ((size_t)q0 * 2 + sp) * hh, *V0 = p.Vf + ((size_t)q0 * 2 + sp) * hh;
  const double *U1 = p.Uf + ((size_t)q1 * 2 + sp) * hh, *V1 = p.Vf + ((size_t)q1 * 2 + sp) * hh;
  const double *Z1 = p.Z + (size_t)q1 * n * n + (size_t)sp * h * (n + 1);  // species block, ld n
  // R_ab = e^{-i alpha m_a} d_ab(beta) e^{-i gamma m_b}: the two phases
  // tabled per state (2h sines and cosines instead of one per packed entry)
  __shared__ double2 ea_s[HMAX], eg_s[HMAX];
  for (int i = tid; i < h; i += NTS) {
    m_off[i] = (short)p.blk_off[i];
    m_sz[i] = (short)p.blk_sz[i];
    m_base[i] = (short)p.blk_base[i];
    ea_s[i] = cexpi(-alp * 0.5 * p.twom[i]);
    eg_s[i] = cexpi(-gam * 0.5 * p.twom[i]);
  }
  __syncthreads();
  for (int e = tid; e < p.npack; e += NTS) {
    const int a = p.pk_a[e], b = p.pk_b[e];
    const double d = p.rtab[(size_t)ib * p.npack + e];
    sR[e] = cscale(d, cmul(ea_s[a], eg_s[b]));
  }
  __syncthreads();
  double2 *RV1 = sm, *RcU1 = sm + hh, *RZ1 = sm + 2 * hh;
  double2 *out = p.sbuf + (((size_t)ip * p.NE + iang) * 2 + sp) * 5 * hh;
  for (int e = tid; e < hh; e += NTS) {
    const int i = e % h, j = e / h;
    const int off = m_off[i], sz = m_sz[i], base = m_base[i];
    double2 a = cz(), b = cz(), c = cz();
    for (int k = 0; k < sz; ++k) {
      const double2 r = sR[base + (i - off) + k * sz];
      rfma(a, V1[(off + k) + j * h], r);
      rfma(b, U1[(off + k) + j * h], cconj(r));
      rfma(c, Z1[(off + k) + (size_t)j * n], r);
    }
    RV1[e] = a;
    RcU1[e] = b;
    RZ1[e] = c;
    out[2 * hh + e] = a;
    out[3 * hh + e] = b;
  }
  __syncthreads();
  // X = V0s^T RV1, Y = U0s^T RcU1, M0 = (R Z1s) R^T
  for (int e = tid; e < hh; e += NTS) {
    const int i = e % h, j = e / h;
    double2 x = cz(), y = cz();
    // (unrolled by four: four terms' loads in flight)
#pragma unroll 4
    for (int k = 0; k < h; ++k) {
      rfma(x, V0[k + i * h], RV1[k + j * h]);
      rfma(y, U0[k + i * h], RcU1[k + j * h]);
    }
    const int off = m_off[j], sz = m_sz[j], base = m_base[j];
    double2 m = cz();
    for (int k = 0; k < sz; ++k) cfma(m, RZ1[i + (off + k) * h], sR[base + (j - off) + k * sz]);
    out[e] = x;
    out[hh + e] = y;
    out[4 * hh + e] = m;
  }
}

struct PhiParams {
  int n, nphi, NE, nwg;
  double cscale;
  const double2 *sbuf;              // k_fac_setup output
  const double *UfT, *VfT;          // [nstates][2][h*h], species blocks, row-major (U_s(j, k) at j h + k)
  const double *Z;                  // [nstates][n*n]
  const double *binv, *pfb;         // Schur route: (-Z_s)^-1 [nstates][2][h*h], Pf(-cscale Z_s) [nstates][2]
  const int *pfs_ok;                // both Pfaffian kernels launched: [nstates] route of bra state q0, else NULL
  const double *t1b;                // n x n
  const double *mtab;               // [3][h]: m, <m+1|J_+|m> (0 at m = j), j(j+1) per species-local state
  const double2 *jtab;              // [2][h^2][4]: coefficients of rho(a,b) in Ekin and the <J^2> pieces (see pnamp_create)
  // species lists (p, n), one term per entry, nl entries padded with zero
  // terms to a multiple of 256: entry q is lofs[q] = LDS byte offsets of
  // rho(k,i), rho(l,j), rho(l,i), rho(k,j) and lofs[nl + q] = those of
  // A01(i,j), A10(k,l)/4 and the matrix element (low, high word); full
  // 32-bit offsets, so the loop unpacks nothing
  const uint4 *lofs[2];
  int nl[2];
  const int *q0, *q1;
  double2 *out;                     // [npairs][NE][2][nphi][FAC_OUT]
  double2 *rho;                     // [npairs][NE][2][nphi][h*h], column-major
  int *status;                      // [npairs][NE][2][nphi]
  // mixed forms: GEMM operands (fragment-ordered, see put_b) per species and
  // the FP64 A01 = kappa01 - kappa01^T dot vectors [npairs][NE][2][nphi][h*h]
  void *brho[2], *brho_lo[2], *ba10[2], *ba10_lo[2];
  double2 *a01;
};

// Fragment order of the V-contraction operands (k_vgemm): element (k, col)
// of a K x ncols operand, 32 columns per tile.  FP32 (v_mfma_f32_32x32x2_f32):
// per k-octet, lane l = (col % 32) + 32 (k % 2) holds k = 8 kq + 2 e + (k % 2),
// e = 0..3, as one 16-byte load.  BF16 (v_mfma_f32_32x32x16_bf16): per
// 16-k block, lane l = (col % 32) + 32 ((k / 8) % 2) holds 8 consecutive k.
template <int KP>
__device__ __forceinline__ size_t swz_f32(int k, size_t col) {
  return ((((col >> 5) * (KP / 8) + (size_t)(k >> 3)) * 64 + (col & 31) + 32 * (k & 1)) << 2) + ((k >> 1) & 3);
}
template <int KP>
__device__ __forceinline__ size_t swz_bf16(int k, size_t col) {
  return ((((col >> 5) * (KP / 16) + (size_t)(k >> 4)) * 64 + (col & 31) + 32 * ((k >> 3) & 1)) << 3) + (k & 7);
}

// bfloat16 by round-to-nearest-even of the FP32 bits (finite inputs)
__host__ __device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_float(unsigned short b) {
  const unsigned u = (unsigned)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// One operand element: FP32, or split BF16 hi + lo with lo = bf16(v - hi).
template <int MIX, int KP>
__device__ __forceinline__ void put_b(void *hi, void *lo, int k, size_t col, double v) {
  if (MIX == MIX_FP32) {
    ((float *)hi)[swz_f32<KP>(k, col)] = (float)v;
  } else {
    const unsigned short bh = bf16_rne((float)v);
    const size_t i = swz_bf16<KP>(k, col);
    ((unsigned short *)hi)[i] = bh;
    ((unsigned short *)lo)[i] = bf16_rne((float)(v - (double)bf16_float(bh)));
  }
}

template <int CTRL, int RM>
__device__ __forceinline__ double dpp_f64(double v) {  // DPP move of a double (0 where no source)
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum of v over the wavefront by the DPP steps of wave_umax, wave-uniform.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_f64<0xB1, 0xF>(v);
  v += dpp_f64<0x4E, 0xF>(v);
  v += dpp_f64<0x141, 0xF>(v);
  v += dpp_f64<0x140, 0xF>(v);
  v += dpp_f64<0x142, 0xA>(v);
  v += dpp_f64<0x143, 0xC>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Gauss-Jordan on an H-row system held one column per lane of the wavefront:
// col[r] is row r of this lane's column, lanes 0..H-1 hold the coefficient
// matrix M, the other lanes right-hand sides.  Step k pivots on row k and the
// unused coefficient column q with the largest |M(k, q)| (|re| + |im| to
// 2^-15 relative, ties to the smallest q): for M = T^T this is the row
// pivoting of T by ZGETRF's IZAMAX.  The row index is static, so nothing is picked out of registers by
// a run-time index; column q reaches every lane by v_readlane from lane q (a
// wave-uniform operand of the update; through LDS the 64-lane broadcast reads
// made the step LDS-bound).  On return the unknown q_k = qmap[k] is row k of
// each right-hand-side column; INFO as LAPACK for an exactly-zero pivot.  The
// workgroup is this one wavefront, so __syncthreads() only orders LDS.
template <int H>
__device__ int wave_gauss_jordan(double2 (&col)[H], int *qmap) {
  const int lane = threadIdx.x;
  unsigned used = 0u;
  int info = 0;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    unsigned key = 0u;
    if (lane < H && !((used >> lane) & 1u)) key = pivot_key_hi<5>(cabs1(col[k]), lane);
    key = wave_umax(key);
    const int q = 31 - (int)(key & 31u);
    used |= 1u << q;
    if (lane == 0) qmap[k] = q;
    auto rl = [&](double x) {
      const long long bb = __double_as_longlong(x);
      return __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(bb >> 32), q) << 32) |
                                  (long long)(unsigned)__builtin_amdgcn_readlane((int)bb, q));
    };
    const double2 pv = make_double2(rl(col[k].x), rl(col[k].y));
    // tested here (uniform) rather than at the INFO store: a test sunk to
    // the use keeps every pivot alive in registers
    if (__builtin_amdgcn_readfirstlane((pv.x == 0.0 && pv.y == 0.0) ? 1 : 0) && info == 0) info = k + 1;
    const double2 prow = cmul(col[k], crecip_nr(pv));
#pragma unroll
    for (int r = 0; r < H; ++r)
      if (r != k) cfms(col[r], make_double2(rl(col[r].x), rl(col[r].y)), prow);
    col[k] = prow;
  }
  __syncthreads();
  return info;
}

// One wavefront per (pair, Euler point, species, gauge angle phi).
//
// Pfaffian of S(phi) (2H x 2H skew) by Parlett-Reid elimination with partial
// pivoting: at step k the pivot of row r = 2k is the largest |S(r, c)|, c > r,
// index c is exchanged with r + 1 (one transposition, sign -1), and the rank-2
// Schur update removes r, r + 1.  ZPfaffianF (Codes/PNAMP/PNAMP_Proj.f90:218-350)
// pivots on the largest entry of the whole remaining matrix; the Pfaffian is
// the same polynomial, the rounding differs.  Its zero rule is kept: a pivot
// of modulus <= 1e-13 gives Pf = 0.  The matrix is held in registers in a 2D
// block-cyclic layout (lane 8a + b holds S(a + 8p, b + 8q), slot rows p <=
// slot columns q); rows r, r + 1 and row / column c of the step go through
// LDS, and the update covers the slot rows that still hold trailing entries.
//
// T = e^{i phi} X + e^{-i phi} Y, W1 = e^{i phi} R V1s T^-1 and
// W2 = e^{-i phi} R* U1s T^-1 solve T^T [W1 W2]^T = [e^{i phi} RV1, e^{-i phi} RcU1]^T
// (Gauss-Jordan on lanes: T^T in lanes 0..H-1, right-hand sides in lanes
// H..3H-1, or in two passes when 3H > 64), then rho = W1 V0s^T,
// kappa10 = W1 U0s^T, kappa01 = -W2 V0s^T (ROKAPPAphiomega,
// Codes/PNAMP/PNAMP_Proj.f90:459-551) with V0s, U0s read as wave-uniform
// operands.
//
// <J^2> pieces (ZJ2_0, Codes/PNAMP/PNAMP_ME2B.f90:1-152) from the ladder form
// sum_k J_k (x) J_k = J_z (x) J_z + (J_+ (x) J_- + J_- (x) J_+) / 2 with J_z
// diagonal and J_+ one off-diagonal inside each orbit:
//   A = sum_a j_a(j_a+1) rho_aa - sum_k Tr(J_k rho J_k rho) + sum_k Tr(J_k kappa10 J_k^T kappa01^T),
//   t_+ = Tr(J_+ rho), t_- = Tr(J_- rho), t_z = Tr(J_z rho),
// so <J^2> = A_p + A_n + (t_zp + t_zn)^2 + (t_+p + t_+n)(t_-p + t_-n).
//
// The eight-term list sum of HFB_Energy_COMPLEX
// (Codes/PNAMP/PNAMP_Energy_OpenMP.f90:5-100) is regrouped exactly: with
// A01 = kappa01 - kappa01^T and A10 = kappa10 - kappa10^T,
//   (1/4) V [r_ijkl - r_jikl - r_ijlk + r_jilk + d (r_klij - r_klji - r_lkij + r_lkji)]
//   = V [rho_ki rho_lj - rho_li rho_kj + A01_ij A10_kl / 4
//        + d (rho_ik rho_jl - rho_jk rho_il + A01_kl A10_ij / 4)].
//
// The work is split in two launches over the same grid: PART_PF the Pfaffian
// (a few KB of LDS and ~100 VGPRs, so many waves per SIMD hide its serial
// steps), PART_REST the rest (13 KB of LDS for rho and one kappa / A block at
// h = 20, so three waves per SIMD).
constexpr int PART_PF = 1, PART_REST = 2;
// PART_PFS: the Pfaffian by the Schur complement of the state block,
//   Pf([[A, -1], [1, B]]) = Pf(B) Pf(A + B^-1),  A = e^{2i phi} M0, B = -Z_s,
// an h x h Pfaffian per (pair, Euler point, species, phi) with B^-1 and
// Pf(cscale B) computed once per state (for states whose Z_s is well
// conditioned; the others take PART_PF)
constexpr int PART_PFS = 3;

// VGPRs of the Pfaffian part when it holds the step's column values in
// registers: the T (T + 1) / 2 complex slots of the 2D layout (T = ceil(2h /
// 8)), 2T held column values, the row pair and ~28 for the rest.
constexpr int pf_hold_vgprs(int h) {
  return 4 * (((2 * h + 7) / 8) * ((2 * h + 7) / 8 + 1) / 2) + 8 * ((2 * h + 7) / 8) + 8 + 20;
}
// Waves per SIMD the register budget is fitted to (the autotuner's knob;
// these are the defaults): the Pfaffian part four where the held column
// values fit 128 VGPRs (h <= 20), else three; the rest three while one
// Gauss-Jordan pass fits the wavefront, else two.
constexpr int pf_default_waves(int h) { return pf_hold_vgprs(h) <= 128 ? 4 : 3; }
template <int H, int PART>
constexpr int fac_default_waves() {
  return PART == PART_PF ? pf_default_waves(H) : PART == PART_PFS ? pf_default_waves(H / 2) : (3 * H <= 64 ? 3 : 2);
}

template <int H, int MIX, int PART, int WV = fac_default_waves<H, PART>()>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WV))) void k_fac_phi(PhiParams p) {
  constexpr int N2 = 2 * H, NU = N2 * (N2 - 1) / 2, SL = (NU + 63) / 64, HH = H * H;
  constexpr int KP = (HH + 31) / 32 * 32;  // V-contraction K padded to the MFMA tile
  constexpr int PASSES = (3 * H <= 64) ? 1 : 2;
  constexpr int HP = H + 1, MS = H * HP;  // LDS leading dimension h + 1: transposed reads avoid bank conflicts
  // B0 = sA[0, MS): rho (ld h + 1; transposed reads avoid bank conflicts);
  // B1 = sA[MS, MS + h^2) (ld h): Gauss-Jordan unknown map (one pass),
  // kappa01, kappa10, then the A-block A01 (strict upper) and A10 / 4
  // (strict lower triangle); 13 KB at h = 20, twelve waves per CU.  The
  // closing reduction reuses sA for 16 quad sums of 14 values (112 double2),
  // more than MS + h^2 below h = 8: LDS stores past the allocation are
  // dropped, so sA is never smaller than that.
  constexpr int RED2 = 16 * 14 / 2;
  __shared__ double2 sA[(MS + HH > RED2) ? MS + HH : RED2];
  // in two passes kappa01 of pass 0 holds B1 through pass 1
  constexpr bool GJ_IN_SA = (PASSES == 1);
  __shared__ int qmap_sep[GJ_IN_SA ? 1 : H];
  double2 *const B0 = sA, *const B1 = sA + MS;
  int *qmap = GJ_IN_SA ? (int *)B1 : qmap_sep;
  const int lane = threadIdx.x;
  // XCD-aware order: consecutive logical indices (the nphi waves of one setup
  // block) stay on one XCD and its L2; gridDim.x is a multiple of 8.
  const unsigned L = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (L >= (unsigned)p.nwg) return;
  const int nphi = p.nphi, n = p.n;
  const int f = (int)(L % (unsigned)nphi);
  unsigned t = L / (unsigned)nphi;
  const int sp = (int)(t & 1u);
  t >>= 1;
  const int iang = (int)(t % (unsigned)p.NE), ip = (int)(t / (unsigned)p.NE);
  const size_t g = ((size_t)ip * p.NE + iang) * 2 + sp;
  const size_t obase = g * nphi + f;
  const double2 *X = p.sbuf + g * 5 * HH, *Y = X + HH, *RV1 = X + 2 * HH, *RcU1 = X + 3 * HH, *M0 = X + 4 * HH;
  const int q0 = p.q0[ip];
  const auto V0T = ro_const(p.VfT + ((size_t)q0 * 2 + sp) * HH), U0T = ro_const(p.UfT + ((size_t)q0 * 2 + sp) * HH);
  const double *Z0 = p.Z + (size_t)q0 * n * n + (size_t)sp * H * (n + 1);  // species block, ld n
  const double phi = (nphi == 1) ? 0.0 : (2.0 * M_PI / nphi) * f;
  const double2 e1 = cexpi(phi), e2 = cexpi(2.0 * phi);

  if constexpr (PART == PART_PF || PART == PART_PFS) {
  // ---- Pfaffian (PART_PF: of the 2h x 2h S; PART_PFS: of A + B^-1, h x h)
  // With both kernels launched over the grid, each keeps the pairs whose bra
  // state takes its route (one pair per wave: the test is wave-uniform)
  if (p.pfs_ok && ((p.pfs_ok[q0] != 0) != (PART == PART_PFS))) return;
  // 2D block-cyclic layout: lane = 8 a + b holds S(a + 8 p, b + 8 q) for
  // 0 <= p <= q < T in slot (p, q); the diagonal slots hold the whole 8 x 8
  // block (both triangles), so every row and every column of the matrix is
  // T slots of one lane octet, and a lane's row and column indices are its
  // own constants plus multiples of 8 (LDS addresses with immediate offsets,
  // no per-entry index tables).  Step k eliminates r = 2k, r + 1 (slot row
  // P = k / 4, a compile-time constant of the unrolled block); the pivot
  // column c (runtime) selects one slot column Q = c / 8 by a uniform switch.
  // Entries that left the trailing matrix keep being updated with values
  // nobody reads again (cheaper than masking them).
  constexpr int MN = (PART == PART_PFS) ? H : N2, MH = MN / 2;  // matrix size, elimination pairs
  constexpr int T = (MN + 7) / 8, NS = T * (T + 1) / 2;
  // Holding the step's column values in registers halves the LDS reads of
  // the update; where they do not fit the variant's register budget (waves
  // per SIMD WV) they are re-read per slot.
  constexpr bool HOLD = pf_hold_vgprs(MH) <= 512 / WV;
  const int a = lane >> 3, b = lane & 7;
  auto SI = [](int pp, int qq) { return pp * T - pp * (pp - 1) / 2 + (qq - pp); };
  // LDS vectors of the step, 8T entries each: R = row r, A = row r + 1,
  // C = column c up to row 8Q + 7 (then the scaled column of the update),
  // D = row c from column 8Q + 8
  __shared__ double2 vec[32 * T];
  double2 *const R = vec, *const A = vec + 8 * T, *const C = vec + 16 * T, *const D = vec + 24 * T;
  double2 sv[NS];
  {
    // S(i, j) for the slot's (i, j): every load unconditional (indices
    // clamped into the blocks), the block chosen by selects
#pragma unroll
    for (int pp = 0; pp < T; ++pp)
#pragma unroll
      for (int qq = pp; qq < T; ++qq) {
        const int i = a + 8 * pp, j = b + 8 * qq, lo = min(i, j), hi = max(i, j);
        const double2 m0v = M0[min(lo, H - 1) + min(hi, H - 1) * H];
        double2 up;  // S(lo, hi)
        if constexpr (PART == PART_PFS) {
          const double bv = p.binv[((size_t)q0 * 2 + sp) * HH + min(lo, H - 1) + min(hi, H - 1) * H];
          up = cmul(e2, m0v);
          up.x += bv;
        } else {
          const double z0v = Z0[(size_t)min(max(lo - H, 0), H - 1) + (size_t)min(max(hi - H, 0), H - 1) * n];
          if (hi < H) up = cmul(e2, m0v);
          else if (lo < H) up = make_double2((hi - H == lo) ? -1.0 : 0.0, 0.0);
          else up = make_double2(-z0v, 0.0);
        }
        up = (hi < MN && lo != hi) ? cscale(p.cscale, up) : cz();
        sv[SI(pp, qq)] = (i < j) ? up : cneg(up);
      }
  }
  double2 pf = make_double2(1.0, 0.0);
  bool zero = false;
#pragma unroll
  for (int P = 0; P < T; ++P) {
    if (zero || 4 * P >= MH - 1) break;
    for (int t = 0; t < 4 && 4 * P + t < MH - 1; ++t) {
      const int r = 8 * P + 2 * t;
      // pivot: the largest |S(r, j)|, j > r (row r is slot row P of lanes
      // a = 2t, the eight lanes 16t .. 16t + 7: reduced among themselves)
      unsigned key = 0u;
      if (a == 2 * t) {
#pragma unroll
        for (int q = P; q < T; ++q) {
          const int j = b + 8 * q;
          const double2 v = sv[SI(P, q)];
          const unsigned kk = pivot_key_hi<6>(v.x * v.x + v.y * v.y, j);
          key = (j > r && j < MN) ? max(key, kk) : key;
        }
      }
      key = (unsigned)__builtin_amdgcn_readlane((int)oct_umax(key), 16 * t);
      const int c = 63 - (int)(key & 63u);
      const bool swap = c != r + 1;
      const int Q = c >> 3, c8 = c & 7;
      // rows r and r + 1 -> R, A; column c -> C (rows below 8Q + 8), row c -> D
      if ((a >> 1) == t) {
        double2 *const dst = ((a & 1) ? A : R) + b;
#pragma unroll
        for (int q = P; q < T; ++q) dst[8 * q] = sv[SI(P, q)];
      }
#pragma unroll
      for (int qq = P; qq < T; ++qq) {
        if (qq == Q) {
          if (b == c8) {
#pragma unroll
            for (int pp = P; pp <= qq; ++pp) C[a + 8 * pp] = sv[SI(pp, qq)];
          }
          if (a == c8) {
#pragma unroll
            for (int q = qq + 1; q < T; ++q) D[b + 8 * q] = sv[SI(qq, q)];
          }
        }
      }
      __syncthreads();
      // pivot S(r, c), wave-uniform
      const double2 sl = R[c];
      const double2 s = make_double2(
          __longlong_as_double(((long long)__builtin_amdgcn_readfirstlane((int)(__double_as_longlong(sl.x) >> 32)) << 32) |
                               (unsigned)__builtin_amdgcn_readfirstlane((int)__double_as_longlong(sl.x))),
          __longlong_as_double(((long long)__builtin_amdgcn_readfirstlane((int)(__double_as_longlong(sl.y) >> 32)) << 32) |
                               (unsigned)__builtin_amdgcn_readfirstlane((int)__double_as_longlong(sl.y))));
      if (!(s.x * s.x + s.y * s.y > 1.0e-26)) {  // |pivot| <= 1e-13
        zero = true;
        break;
      }
      const double2 inv = crecip_nr(s);
      pf = cmul(pf, swap ? cneg(s) : s);
      // The exchange c <-> r + 1 of the trailing indices: position c takes
      // index r + 1, so S'(x, c) = S(x, r + 1), S'(c, y) = S(r + 1, y) (= -A, A)
      if (swap) {
#pragma unroll
        for (int qq = P; qq < T; ++qq) {
          if (qq == Q) {
            if (b == c8) {
#pragma unroll
              for (int pp = P; pp <= qq; ++pp) sv[SI(pp, qq)] = cneg(A[a + 8 * pp]);
            }
            if (a == c8) {
#pragma unroll
              for (int q = qq; q < T; ++q) sv[SI(qq, q)] = A[b + 8 * q];
            }
          }
        }
      }
      // Schur update with the columns of the exchanged matrix,
      //   S'(i, j) = S(i, j) + S(i, r+1) S(j, r) / s - S(i, r) S(j, r+1) / s,
      // S(x, r) = -R[x] (R[c] <- R[r + 1]), S(x, r+1) / s -> C[x] (C[c] <- S(r + 1, c) / s)
      if (lane >= r + 2 && lane < MN) {
        const bool xc = swap && lane == c;
        const bool low = lane < 8 * Q + 8;
        const double2 v = xc ? C[r + 1] : (low ? C[lane] : D[lane]);
        C[lane] = cmul(v, (xc || low) ? inv : cneg(inv));
        if (xc) R[c] = R[r + 1];
      }
      __syncthreads();
      if constexpr (HOLD) {
        double2 Rj[T], Cj[T];
#pragma unroll
        for (int q = P; q < T; ++q) {
          Rj[q] = R[b + 8 * q];
          Cj[q] = C[b + 8 * q];
        }
#pragma unroll
        for (int pp = P; pp < T; ++pp) {
          const double2 Ri = R[a + 8 * pp], Ci = C[a + 8 * pp];
#pragma unroll
          for (int q = pp; q < T; ++q) {
            cfms(sv[SI(pp, q)], Ci, Rj[q]);  // S += C_i (-R_j) - (-R_i) C_j
            cfma(sv[SI(pp, q)], Ri, Cj[q]);
          }
        }
      } else {
#pragma unroll
        for (int pp = P; pp < T; ++pp) {
          const double2 Ri = R[a + 8 * pp], Ci = C[a + 8 * pp];
#pragma unroll
          for (int q = pp; q < T; ++q) {
            const double2 Rj = R[b + 8 * q], Cj = C[b + 8 * q];
            cfms(sv[SI(pp, q)], Ci, Rj);
            cfma(sv[SI(pp, q)], Ri, Cj);
          }
        }
      }
      __syncthreads();
    }
  }
  if (zero) {
    pf = cz();
  } else {
    // the last 2 x 2 block: S(MN - 2, MN - 1), one lane of slot (T - 1, T - 1)
    constexpr int LL = ((MN - 2) & 7) * 8 + ((MN - 1) & 7);
    const double2 v = sv[SI((MN - 2) >> 3, (MN - 1) >> 3)];
    pf = cmul(pf, make_double2(__shfl(v.x, LL), __shfl(v.y, LL)));
  }

  if constexpr (PART == PART_PFS) pf = cscale(p.pfb[(size_t)q0 * 2 + sp], pf);  // Pf(cscale B)
  if (lane == 0) p.out[obase * FAC_OUT] = pf;
  } else {
  // ---- Gauss-Jordan and the density products
  // One pass: T^T in lanes 0..H-1, RV1 (-> W1) in H..2H-1, RcU1 (-> W2) in
  // 2H..3H-1.  Two passes: pass 0 RcU1 (-> W2), pass 1 RV1 (-> W1), both in
  // lanes H..2H-1.  W1 rows give rho (B0) and kappa10, W2 rows kappa01 (B1).
  int info = 0;
  constexpr int K01L = (PASSES == 1) ? 2 * H : H;  // first lane of the W2 rows
  const bool w1lane = lane >= H && lane < 2 * H;
  const int irow = (lane >= K01L && lane < K01L + H && PASSES == 1) ? lane - K01L : lane - H;
  double2 *rho_out = p.rho + obase * HH;
  const size_t mcol = ((size_t)ip * p.NE + iang) * nphi + f;  // GEMM column pair of this phi in species sp
  constexpr int NE64 = (HH + 63) / 64;
  double2 a01r[NE64];  // A01 = kappa01 - kappa01^T, strict upper triangle
  double2 ajk = cz();  // kappa part of <J^2>
  const auto mz = ro_const(p.mtab), cp = ro_const(p.mtab + H);
#pragma unroll
  for (int pass = 0; pass < PASSES; ++pass) {
    const bool isW1 = w1lane && (PASSES == 1 || pass == 1);
    const bool isW2 = (PASSES == 1) ? (lane >= K01L && lane < K01L + H) : (pass == 0 && w1lane);
    double2 col[H];
    if (lane < H) {
#pragma unroll
      for (int r = 0; r < H; ++r)
        col[r] = cadd(cmul(e1, X[lane + r * H]), cmul(cconj(e1), Y[lane + r * H]));
    } else if (isW1 || isW2) {
      const double2 *src = isW2 ? RcU1 : RV1;
      const double2 ph = isW2 ? cconj(e1) : e1;
#pragma unroll
      for (int r = 0; r < H; ++r) col[r] = cmul(ph, src[irow + r * H]);
    } else {
#pragma unroll
      for (int r = 0; r < H; ++r) col[r] = cz();
    }
    const int inf = wave_gauss_jordan<H>(col, qmap);
    if (inf && !info) info = inf;
    // solution rows in unknown order through B0, h rows at a time (unknown
    // qmap[k] is row k), back into col: w[k] is W(irow, k) on the W lanes
#pragma unroll
    for (int g = 0; g < ((PASSES == 1) ? 2 : 1); ++g) {
      const bool gl = lane >= H * (g + 1) && lane < H * (g + 2);
      if (gl) {
#pragma unroll
        for (int r = 0; r < H; ++r) B0[(lane - H * (g + 1)) * HP + qmap[r]] = col[r];
      }
      __syncthreads();
      if (gl) {
#pragma unroll
        for (int k = 0; k < H; ++k) col[k] = B0[(lane - H * (g + 1)) * HP + k];
      }
      __syncthreads();
    }
    const double2 (&w)[H] = col;
    // V0s sweep: rho (W1 rows) -> B0 and HBM, kappa01 (W2 rows) -> B1
#pragma unroll 2
    for (int j = 0; j < H; ++j) {
      double2 acc = cz();
#pragma unroll
      for (int k = 0; k < H; ++k) rfma(acc, V0T[j * H + k], w[k]);
      if (isW1) {
        B0[irow + j * HP] = acc;
        rho_out[irow + j * H] = acc;
        if (MIX != MIX_FP64) {
          put_b<MIX, KP>(p.brho[sp], p.brho_lo[sp], irow + j * H, 2 * mcol, acc.x);
          put_b<MIX, KP>(p.brho[sp], p.brho_lo[sp], irow + j * H, 2 * mcol + 1, acc.y);
        }
      }
      if (isW2) B1[irow + j * H] = cneg(acc);
    }
    __syncthreads();
    if (PASSES == 1 || pass == 1) {
#pragma unroll
      for (int u = 0; u < NE64; ++u) {
        const int e = lane + 64 * u, a = e % H, b = e / H;
        a01r[u] = (e < HH && a < b) ? csub(B1[a + b * H], B1[b + a * H]) : cz();
      }
      // U0s sweep: kappa10 (W1 rows) with the kappa part of <J^2>,
      //   sum_ab kappa10(a,b) [m_a m_b kappa01(a,b) + c_a c_{b+1} kappa01(a-1,b+1) / 2
      //                        + c_{a+1} c_b kappa01(a+1,b-1) / 2]
      // (c_a = <m_a + 1|J_+|m_a>, 0 at m_a = j_a, so a - 1 is in a's orbit
      // when c_a != 0); kappa10(a, b - 1) replaces kappa01(a, b - 1) in B1 one
      // column late, after the last read of that column
      const int a = w1lane ? irow : 0;
      const double ma = mz[a], ca = cp[a], ca1 = (a + 1 < H) ? cp[a + 1] : 0.0;
      const int am = (a > 0) ? a - 1 : 0, ap = (a + 1 < H) ? a + 1 : a;
      double2 prev = cz();
#pragma unroll 2
      for (int j = 0; j < H; ++j) {
        double2 acc = cz();
#pragma unroll
        for (int k = 0; k < H; ++k) rfma(acc, U0T[j * H + k], w[k]);
        if (w1lane) {
          double2 x = cscale(ma * mz[j], B1[a + j * H]);
          if (j + 1 < H) x = cadd(x, cscale(0.5 * ca * cp[j + 1], B1[am + (j + 1) * H]));
          if (j > 0) x = cadd(x, cscale(0.5 * ca1 * cp[j], B1[ap + (j - 1) * H]));
          cfma(ajk, acc, x);
          if (j > 0) B1[a + (j - 1) * H] = prev;
          prev = acc;
        }
      }
      if (w1lane) B1[a + (H - 1) * H] = prev;
      __syncthreads();
    }
  }

  // ---- traces, the rho part of <J^2>, kinetic energy
  // per element rho(a, b): Ekin t1(b, a); <J^2> -m_a m_b rho(b, a) - c_a c_b
  // rho(b-1, a-1) + j_a(j_a+1) d_ab; traces d_ab, m_a d_ab, c_a d_{a,b+1},
  // c_b d_{b,a+1} -- eight coefficients tabled per element, no branches
  double2 tr = cz(), ek = cz(), aj = ajk, tp = cz(), tm = cz(), tz = cz();
  const double2 *jt = p.jtab + (size_t)sp * HH * 4;
  for (int e = lane; e < HH; e += 64) {
    const int a = e % H, b = e / H;
    const double2 c01 = jt[4 * e], c23 = jt[4 * e + 1], c45 = jt[4 * e + 2], c67 = jt[4 * e + 3];
    const double2 rab = B0[a + b * HP];
    const double2 rba = B0[b + a * HP];
    const double2 rl = B0[(b > 0 ? b - 1 : 0) + (a > 0 ? a - 1 : 0) * HP];
    rfma(ek, c01.x, rab);
    rfma(aj, -c01.y, cmul(rab, rba));
    rfma(aj, -c23.x, cmul(rab, rl));
    rfma(tr, c23.y, rab);
    rfma(tz, c45.x, rab);
    rfma(aj, c45.y, rab);
    rfma(tp, c67.x, rab);
    rfma(tm, c67.y, rab);
  }
  // A10 / 4 = (kappa10 - kappa10^T) / 4, then the A-block
  double2 a10r[NE64];
#pragma unroll
  for (int u = 0; u < NE64; ++u) {
    const int e = lane + 64 * u, a = e % H, b = e / H;
    a10r[u] = (e < HH && a < b) ? cscale(0.25, csub(B1[a + b * H], B1[b + a * H])) : cz();
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NE64; ++u) {
    const int e = lane + 64 * u, a = e % H, b = e / H;
    if (e < HH && a < b) {
      B1[a + b * H] = a01r[u];
      B1[b + a * H] = a10r[u];
    }
  }
  __syncthreads();
  double2 e2b = cz();
  if (MIX != MIX_FP64) {
    // the species list moves to k_vgemm: A01 (dot vector) and A10 (operand)
    double2 *a01 = p.a01 + obase * HH;
    for (int e = lane; e < HH; e += 64) {
      const int a = e % H, b = e / H;
      const double2 x01 = (a < b) ? B1[a + b * H] : (a > b ? cneg(B1[b + a * H]) : cz());
      const double2 x10 = (a < b) ? cscale(4.0, B1[b + a * H]) : (a > b ? cscale(-4.0, B1[a + b * H]) : cz());
      a01[e] = x01;
      put_b<MIX, KP>(p.ba10[sp], p.ba10_lo[sp], e, 2 * mcol, x10.x);
      put_b<MIX, KP>(p.ba10[sp], p.ba10_lo[sp], e, 2 * mcol + 1, x10.y);
    }
  } else {
    // ---- species list: V [rho_ki rho_lj - rho_li rho_kj + A01_ij A10_kl / 4]
    const uint4 *lo = p.lofs[sp];
    const int nl = p.nl[sp];
    const char *lds = (const char *)sA;
    auto ld = [&](unsigned off) { return *(const double2 *)(lds + off); };
    for (int q0 = lane; q0 < nl; q0 += 4 * 64) {
      uint4 o[4], w[4];  // four entries loaded ahead of their LDS reads
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        o[u] = lo[q0 + 64 * u];
        w[u] = lo[nl + q0 + 64 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double2 r1 = ld(o[u].x), r2 = ld(o[u].y), r3 = ld(o[u].z), r4 = ld(o[u].w), x = ld(w[u].x),
                      y = ld(w[u].y);
        const double me = __hiloint2double((int)w[u].w, (int)w[u].z);
        double2 t = cmul(r1, r2);
        cfms(t, r3, r4);
        cfma(t, x, y);
        rfma(e2b, me, t);
      }
    }
  }
  // the 14 sums: quad sums by DPP, the 16 quad sums of each value through
  // LDS (sA is free after the list), lane q < 14 adds those of value q and
  // writes it (Re, Im of output slots 1..7)
  double v[14] = {e2b.x, e2b.y, ek.x, ek.y, tr.x, tr.y, aj.x, aj.y, tp.x, tp.y, tm.x, tm.y, tz.x, tz.y};
#pragma unroll
  for (int q = 0; q < 14; ++q) {
    v[q] += dpp_f64<0xB1, 0xF>(v[q]);  // quad_perm [1,0,3,2]
    v[q] += dpp_f64<0x4E, 0xF>(v[q]);  // quad_perm [2,3,0,1]
  }
  double *red = (double *)sA;
  static_assert(sizeof(sA) >= 16 * 14 * sizeof(double), "sA holds the closing reduction");
  __syncthreads();
  if ((lane & 3) == 0) {
#pragma unroll
    for (int q = 0; q < 14; ++q) red[(lane >> 2) * 14 + q] = v[q];
  }
  __syncthreads();
  double2 *o = p.out + obase * FAC_OUT;
  if (lane < 14) {
    double tot = 0.0;
#pragma unroll
    for (int qd = 0; qd < 16; ++qd) tot += red[qd * 14 + lane];
    ((double *)o)[2 + lane] = tot;
  }
  if (lane == 0) {
    o[8] = cz();
    p.status[obase] = info;
  }
  }
}

struct CombParams {
  int h, NE, zval, nval;
  double vac_scale, snpfaf, sgn_h, e0;
  const double *det;
  const int *q0, *q1;
  const int *pz, *pn;               // valence numbers per pair, or NULL
  // W_pn (rows x = c + a h, columns y = d + b h) in dense blocks of equal
  // Delta m = m_c - m_a = -(m_d - m_b): the interaction conserves M
  const int2 *pn_rows;              // [h^2] by block order: (x, block)
  const int *pn_blk;                // [nblk][5]: first row position, rows, first column (in pn_ycol), columns, offset in pn_w
  const int *pn_ycol;               // [h^2] column lists of the blocks
  const double *pn_w;               // blocks, column-major
  int pn_ncol;                      // length of pn_ycol
  const double2 *out, *rho;
  const int *fstatus;               // [npairs][NE][2][nphi]
  const float2 *gpn;                // mixed forms: G = W_pn rho_n from k_vgemm, [npairs][NE][nphi][h*h]
  double2 *samples;                 // optional: [npairs][NE][nphi][nphi][2] = norm, energy per gauge pair
  double2 *zres;                    // [npairs][NE][NOBS]
  int *status;                      // [npairs][NE]
};

// One workgroup per (pair, Euler point).  The proton-neutron energy is the
// bilinear form E_pn(phi_p, phi_n) = rho_p(phi_p)^T W rho_n(phi_n) with
// W[(c,a),(d,b)] = V_abcd and W[(a,c),(b,d)] += V_abcd for list entries with
// a + b != c + d (for a, c protons and b, d neutrons the eight-term sum of
// HFB_Energy_COMPLEX reduces to V [rho_p(c,a) rho_n(d,b) + rho_p(a,c) rho_n(b,d)]).
// G = W rho_n(phi_n) is formed once per phi_n from the dense Delta-m blocks of W, then
// the nphi^2 dot products with rho_p(phi_p); the gauge points are summed into
// ZRES.  NP = nphi is a template argument so the per-gauge register arrays
// stay in registers.  With GMIX (mixed forms) G comes from k_vgemm instead.
template <int NP, bool GMIX>
// (up to nine gauge angles four waves per SIMD: two eight-wave workgroups per
// CU, as many as their LDS allows; more angles need more than 128 VGPRs)
__global__ __launch_bounds__(NTC) __attribute__((amdgpu_waves_per_eu(NP <= 9 ? 4 : 2))) void k_fac_combine(CombParams p) {
  extern __shared__ double2 dsm[];  // rho_n -> G [NP][h^2], then the block column lists; rho_p from global memory
  __shared__ double2 epn[NP * NP];
  __shared__ double red[(NTC + 63) / 64 * 2 * NOBS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int iang = blockIdx.x, ip = blockIdx.y;
  const int h = p.h, hh = h * h;
  const size_t base = ((size_t)ip * p.NE + iang) * 2;
  double2 *rn = dsm;
  int *ycol = (int *)(dsm + (size_t)NP * hh);
  // rho_n (or G) into LDS: every load of a thread issued before the first
  // store (indices clamped, no load in a branch), so the HBM latency is
  // paid once per thread rather than once per element
  for (int e = tid; e < p.pn_ncol; e += NTC) ycol[e] = p.pn_ycol[e];
  {
    constexpr int NB = 8;  // loads in flight per thread (the register cap is 128)
    const int nt = NP * hh;
    for (int e0 = 0; e0 < nt; e0 += NB * NTC) {
      double2 tmp[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = min(e0 + tid + NTC * u, nt - 1);
        if (GMIX) {
          const float2 gv = p.gpn[((size_t)ip * p.NE + iang) * NP * hh + e];
          tmp[u] = make_double2(gv.x, gv.y);
        } else {
          tmp[u] = p.rho[(base + 1) * NP * hh + e];
        }
      }
      // unconditional stores: a clamped element is element nt - 1, stored
      // again with its own value (a store in a branch would pull its load
      // into the branch and wait on it there)
#pragma unroll
      for (int u = 0; u < NB; ++u) rn[min(e0 + tid + NTC * u, nt - 1)] = tmp[u];
    }
  }
  __syncthreads();
  if (!GMIX) {
  // G rows in block order (row position tid + r NTC), all phi_n, in
  // registers.  The lanes of one block walk the same column list, so each
  // rho_n read is an LDS broadcast.
  constexpr int RPT = (FAC_HMAX * FAC_HMAX + NTC - 1) / NTC;
  double2 g[RPT][NP];
  int xrow[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
#pragma unroll
    for (int f = 0; f < NP; ++f) g[r][f] = cz();
    const int px = tid + r * NTC;
    xrow[r] = -1;
    if (px < hh) {
      const int2 rw = p.pn_rows[px];
      const int *bk = p.pn_blk + 5 * rw.y;
      const int nrows = bk[1], ncol = bk[3];
      const int *yc = ycol + bk[2];
      const double *wb = p.pn_w + bk[4] + (px - bk[0]);
      xrow[r] = rw.x;
      // four block entries per group, the next group's loads issued before
      // this group's products (indices clamped to the last entry, weight 0
      // past it)
      double v0 = wb[0], v1 = wb[(size_t)min(1, ncol - 1) * nrows], v2 = wb[(size_t)min(2, ncol - 1) * nrows],
             v3 = wb[(size_t)min(3, ncol - 1) * nrows];
      for (int k = 0; k < ncol; k += 4) {
        const double c0 = v0, c1 = v1, c2 = v2, c3 = v3;
        v0 = wb[(size_t)min(k + 4, ncol - 1) * nrows];
        v1 = wb[(size_t)min(k + 5, ncol - 1) * nrows];
        v2 = wb[(size_t)min(k + 6, ncol - 1) * nrows];
        v3 = wb[(size_t)min(k + 7, ncol - 1) * nrows];
        const double cv[4] = {c0, c1, c2, c3};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool ok = k + u < ncol;
          const int y = yc[ok ? k + u : 0];
          const double vv = ok ? cv[u] : 0.0;
#pragma unroll
          for (int f = 0; f < NP; ++f) rfma(g[r][f], vv, rn[(size_t)f * hh + y]);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int x = xrow[r];
    if (x >= 0) {
#pragma unroll
      for (int f = 0; f < NP; ++f) rn[(size_t)f * hh + x] = g[r][f];
    }
  }
  __syncthreads();
  }
  const int grp = tid >> 5, gl = tid & 31;
  for (int fp = grp; fp < NP; fp += NTC / 32) {
    double2 acc[NP];
#pragma unroll
    for (int f = 0; f < NP; ++f) acc[f] = cz();
    const double2 *rpf = p.rho + (base * NP + fp) * hh;
    // this lane's rho_p elements x = gl + 32 i, eight loads issued together
    // (clamped; weight 0 past h^2)
    for (int x0 = 0; x0 < hh; x0 += 8 * 32) {
      double2 ra[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) ra[i] = rpf[min(x0 + gl + 32 * i, hh - 1)];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int x = x0 + gl + 32 * i, xc = min(x, hh - 1);
        const double2 av = (x < hh) ? ra[i] : cz();
#pragma unroll
        for (int f = 0; f < NP; ++f) cfma(acc[f], av, rn[(size_t)f * hh + xc]);
      }
    }
#pragma unroll
    for (int f = 0; f < NP; ++f) {
      double re = acc[f].x, im = acc[f].y;
      for (int off = 16; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
      }
      if (gl == 0) epn[fp * NP + f] = make_double2(re, im);
    }
  }
  __syncthreads();
  const double2 *op = p.out + base * NP * FAC_OUT, *on = p.out + (base + 1) * NP * FAC_OUT;
  const double vac = p.snpfaf * p.sgn_h * (p.det[p.q0[ip]] * p.vac_scale) * (p.det[p.q1[ip]] * p.vac_scale);
  double v[2 * NOBS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = tid; t < NP * NP; t += NTC) {
    const int fp = t / NP, fn = t % NP;
    const double phip = (NP == 1) ? 0.0 : (2.0 * M_PI / NP) * fp;
    const double phin = (NP == 1) ? 0.0 : (2.0 * M_PI / NP) * fn;
    const double2 *a = op + fp * FAC_OUT, *b = on + fn * FAC_OUT;
    const double2 zn = cmul(cscale(vac, cmul(a[0], b[0])),
                            cexpi(-phip * (p.pz ? p.pz[ip] : p.zval) - phin * (p.pn ? p.pn[ip] : p.nval)));
    const double2 ekin = cadd(a[2], b[2]);
    const double2 e2s = cadd(cadd(a[1], b[1]), cadd(a[8], b[8]));
    const double2 E = cadd(cadd(e2s, epn[fp * NP + fn]), cadd(ekin, make_double2(p.e0, 0.0)));
    // <J^2> = A_p + A_n + (t_zp + t_zn)^2 + (t_+p + t_+n)(t_-p + t_-n)
    const double2 tzs = cadd(a[7], b[7]);
    const double2 J2 = cadd(cadd(cadd(a[4], b[4]), cmul(tzs, tzs)), cmul(cadd(a[5], b[5]), cadd(a[6], b[6])));
    const double2 o[NOBS] = {zn, cmul(zn, a[3]), cmul(zn, b[3]), cmul(zn, ekin), cmul(zn, E), cmul(zn, J2)};
    if (p.samples) {
      double2 *sm = p.samples + (((size_t)ip * p.NE + iang) * NP * NP + t) * 2;
      sm[0] = zn;
      sm[1] = E;
    }
#pragma unroll
    for (int k = 0; k < NOBS; ++k) {
      v[2 * k] += o[k].x;
      v[2 * k + 1] += o[k].y;
    }
  }
#pragma unroll
  for (int k = 0; k < 2 * NOBS; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  if (lane == 0)
    for (int k = 0; k < 2 * NOBS; ++k) red[wave * 2 * NOBS + k] = v[k];
  __syncthreads();
  if (tid == 0) {
    const double inv = 1.0 / (double)(NP * NP);
    double2 *z = p.zres + ((size_t)ip * p.NE + iang) * NOBS;
    for (int k = 0; k < NOBS; ++k) {
      double re = 0, im = 0;
      for (int w = 0; w < (NTC + 63) / 64; ++w) {
        re += red[w * 2 * NOBS + 2 * k];
        im += red[w * 2 * NOBS + 2 * k + 1];
      }
      z[k] = make_double2(inv * re, inv * im);
    }
    int st = 0;
    for (int q = 0; q < 2 * NP && !st; ++q) st = p.fstatus[base * NP + q];
    p.status[(size_t)ip * p.NE + iang] = st;
  }
}

// E2 observables of the factored form, one workgroup per (pair, Euler point),
// launched after k_fac_combine: the species densities rho_s(phi) that
// k_fac_phi wrote give, per species and gauge angle, tr_s(Q_mu rho_s(phi))
// (ZQ2, Codes/PNAMP/PNAMP_ME1B.f90:52-75) and the orbit-pair traces
// sum_m rho(k_a(m), k_b(m)) (ZRho1B_Angles, Codes/PNAMP/PNAMP_Rho1B.f90:13-33).
// A proton quantity depends on phi_p only, so its gauge sum weighs it with the
// norm summed over phi_n (and a neutron one with the norm summed over phi_p).
struct FacE2Params {
  int n, h, NE, nphi, zval, nval, nob2;
  double vac_scale, snpfaf, sgn_h;
  const double *det;
  const int *q0, *q1;
  const int *pz, *pn;                // valence numbers per pair, or NULL
  const double2 *out, *rho;          // k_fac_phi outputs
  const double *q2;                  // [5][n*n] column-major
  const int4 *orbp;                  // [nob2]: species, first m-state of a, of b (global), size
  double2 *ze2, *zrho;               // [npairs][NE][10], [npairs][NE][nob2]
};

__global__ __launch_bounds__(256) void k_fac_e2(FacE2Params p) {
  __shared__ double2 zn[NPHI_MAX * NPHI_MAX];
  __shared__ double2 wsum[2][NPHI_MAX];        // norm summed over phi_n (per phi_p), over phi_p (per phi_n)
  __shared__ double2 tq[2][NPHI_MAX][5];
  __shared__ double2 td[NPHI_MAX][NOB2_MAX];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int iang = blockIdx.x, ip = blockIdx.y;
  const int h = p.h, hh = h * h, n = p.n, np = p.nphi;
  const size_t base = ((size_t)ip * p.NE + iang) * 2;
  const double vac = p.snpfaf * p.sgn_h * (p.det[p.q0[ip]] * p.vac_scale) * (p.det[p.q1[ip]] * p.vac_scale);
  for (int t = tid; t < np * np; t += blockDim.x) {
    const int fp = t / np, fn = t % np;
    const double phip = (np == 1) ? 0.0 : (2.0 * M_PI / np) * fp;
    const double phin = (np == 1) ? 0.0 : (2.0 * M_PI / np) * fn;
    const double2 a = p.out[(base * np + fp) * FAC_OUT], b = p.out[((base + 1) * np + fn) * FAC_OUT];
    zn[t] = cmul(cscale(vac, cmul(a, b)),
                 cexpi(-phip * (p.pz ? p.pz[ip] : p.zval) - phin * (p.pn ? p.pn[ip] : p.nval)));
  }
  // one wave per (species, phi): tr_s(Q_mu rho_s(phi)) for the five mu
  for (int u = wave; u < 2 * np; u += blockDim.x / 64) {
    const int sp = u / np, f = u % np;
    const double2 *r = p.rho + ((base + sp) * np + f) * hh;
    double q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = lane; e < hh; e += 64) {
      const int a = e % h, b = e / h;
      const double2 rba = r[b + a * h];  // rho(b, a), column-major h x h
      const size_t qi = (size_t)(sp * h + a) + (size_t)(sp * h + b) * n;
      for (int mu = 0; mu < 5; ++mu) {
        const double qv = p.q2[(size_t)mu * n * n + qi];
        q[2 * mu] = fma(qv, rba.x, q[2 * mu]);
        q[2 * mu + 1] = fma(qv, rba.y, q[2 * mu + 1]);
      }
    }
    for (int k = 0; k < 10; ++k) q[k] = wave_sum_dpp(q[k]);
    if (lane < 5) tq[sp][f][lane] = make_double2(q[2 * lane], q[2 * lane + 1]);
  }
  for (int t = tid; t < np * p.nob2; t += blockDim.x) {
    const int f = t / p.nob2, k = t % p.nob2;
    const int4 op = p.orbp[k];
    const double2 *r = p.rho + ((base + op.x) * np + f) * hh;
    const int fa = op.y - op.x * h, fb = op.z - op.x * h;
    double2 acc = cz();
    for (int m = 0; m < op.w; ++m) acc = cadd(acc, r[(fa + m) + (fb + m) * h]);
    td[f][k] = acc;
  }
  __syncthreads();
  if (tid < 2 * np) {
    const int sp = tid / np, f = tid % np;
    double2 acc = cz();
    for (int g = 0; g < np; ++g) acc = cadd(acc, sp == 0 ? zn[f * np + g] : zn[g * np + f]);
    wsum[sp][f] = acc;
  }
  __syncthreads();
  const double inv = 1.0 / (double)(np * np);
  if (tid < 10) {
    const int sp = tid / 5, mu = tid % 5;
    double2 acc = cz();
    for (int f = 0; f < np; ++f) acc = cadd(acc, cmul(wsum[sp][f], tq[sp][f][mu]));
    p.ze2[((size_t)ip * p.NE + iang) * 10 + tid] = cscale(inv, acc);
  }
  for (int k = tid; k < p.nob2; k += blockDim.x) {
    const int sp = p.orbp[k].x;
    double2 acc = cz();
    for (int f = 0; f < np; ++f) acc = cadd(acc, cmul(wsum[sp][f], td[f][k]));
    p.zrho[((size_t)ip * p.NE + iang) * p.nob2 + k] = cscale(inv, acc);
  }
}

// Reduced E2 matrix elements from the projected E2 observables
// I_mu(J2, K', K2) = <Phi0| Q_{2mu} P^{J2}_{K' K2} |Phi1> (the plain
// projection of ze2): qred(s, J1 - J2, J2, K1, K2) = sum_mu T(J1 - J2, J2, K1, mu)
// I_{s,mu}(J2, K1 - mu, K2), T = (2 J1 + 1) (-1)^(J1 - K1) (J1 2 J2; -K1 mu K1 - mu)
// tabled on the host (the sums of Codes/PNAMP/PNAMP_Kernel3D.f90:590-710).
// One workgroup per pair; rho1b: wJ(0) sum_O w zrho(O) / sqrt(2 j_a + 1)
// (ZRho1B_Integral, Codes/PNAMP/PNAMP_Rho1B.f90:39-62).
struct E2RedParams {
  int jmax, NE, nob2, nalp, ngam;
  const double2 *iproj;              // [npairs][10][jmax+1][2jmax+1][2jmax+1]
  const double *t3j;                 // [5][jmax+1][2jmax+5][5]
  double2 *qred;                     // [npairs][2][5][jmax+1][2jmax+5][2jmax+1]
  const double2 *zrho;               // [npairs][NE][nob2]
  const double *walp, *wbet, *wgam;
  const double *orbscale;            // [nob2]: wJ(0) / sqrt(2 j_a + 1)
  double2 *rho1b;                    // [npairs][nob2]
};

__global__ __launch_bounds__(256) void k_e2_reduce(E2RedParams p) {
  const int ip = blockIdx.x, tid = threadIdx.x;
  const int jm = p.jmax, J2m = 2 * jm + 1, K1d = 2 * jm + 5;
  const size_t per_io = (size_t)(jm + 1) * J2m * J2m;
  const double2 *I = p.iproj + (size_t)ip * 10 * per_io;
  const int nq = 2 * 5 * (jm + 1) * K1d * J2m;
  for (int o = tid; o < nq; o += blockDim.x) {
    const int K2 = o % J2m - jm, K1 = (o / J2m) % K1d - jm - 2, J2 = (o / (J2m * K1d)) % (jm + 1);
    const int dJ = (o / (J2m * K1d * (jm + 1))) % 5 - 2, sp = o / (J2m * K1d * (jm + 1) * 5);
    const int J1 = J2 + dJ;
    double2 acc = cz();
    if (J1 >= 0 && K1 >= -J1 && K1 <= J1 && K2 >= -J2 && K2 <= J2) {
      for (int mu = -2; mu <= 2; ++mu) {
        const int Kp = K1 - mu;
        if (Kp < -J2 || Kp > J2) continue;
        const double t = p.t3j[(((size_t)(dJ + 2) * (jm + 1) + J2) * K1d + (K1 + jm + 2)) * 5 + (mu + 2)];
        const double2 v = I[(size_t)(sp * 5 + mu + 2) * per_io + ((size_t)J2 * J2m + (Kp + jm)) * J2m + (K2 + jm)];
        acc = cadd(acc, cscale(t, v));
      }
    }
    p.qred[(size_t)ip * nq + o] = acc;
  }
  for (int k = tid; k < p.nob2; k += blockDim.x) {
    double2 acc = cz();
    for (int i = 0; i < p.NE; ++i) {
      const int ig = i % p.ngam, ia = (i / p.ngam) % p.nalp, ib = i / (p.ngam * p.nalp);
      acc = cadd(acc, cscale(p.walp[ia] * p.wbet[ib] * p.wgam[ig], p.zrho[((size_t)ip * p.NE + i) * p.nob2 + k]));
    }
    p.rho1b[(size_t)ip * p.nob2 + k] = cscale(p.orbscale[k], acc);
  }
}

// ------------------------------------------------------------------ V contraction on MFMA (mixed forms)
// The quadratic forms of the V list as dense h^2 x h^2 matrices applied to
// all density columns of a chunk at once, as in HFB_FIELD_COMPLEX
// (Codes/GCM/VAPHFB_HFB_Field.f90:4-123) with E = Tr(t rho) + Tr(Gamma rho) / 2
// + Tr(Delta10 kappa01^T) / 2 (Codes/PNAMP/PNAMP_HFB_Energy.f90:1-35):
//   E_rho(s)   = rho_s^T (W_ss rho_s),  E_kappa(s) = A01_s^T (Q_ss A10_s),
//   G_pn       = W_pn rho_n, dotted with rho_p(phi_p) by k_fac_combine,
// W and Q built from the species lists so that they reproduce the eight-term
// sum of HFB_Energy_COMPLEX exactly (see pnamp_create).  C = A B with A = W or
// Q (KP x KP, fragment order), B = density columns (Re and Im of each phi as
// two columns).  A workgroup of four waves covers 128 columns, a wave 32
// columns and all KP / 32 row blocks.  The epilogue dots C with the FP64
// vector of its column pair (rho or A01) and writes one complex number per
// phi, or stores G in FP32.  FP32: v_mfma_f32_32x32x2_f32.  BF16: split
// operands x = hi + lo, three products hi*hi + hi*lo + lo*hi per
// v_mfma_f32_32x32x16_bf16, FP32 accumulation.
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct VProblem {
  const void *ahi, *alo, *bhi, *blo;
  const double2 *dot;               // mode 0: [npairs][NE][2][nphi][h*h] (rho or A01)
  float *g;                         // mode 1: [npairs][NE][nphi][h*h] complex FP32
  int mode, species, slot, ncols;
};
constexpr int NVPROB = 5;           // W_pp rho_p, W_nn rho_n, Q_pp A10_p, Q_nn A10_n, W_pn rho_n
struct VGemmParams {
  VProblem prob[NVPROB];
  int h, nphi;
  double2 *out;                     // [npairs][NE][2][nphi][FAC_OUT]
};

// LDS bytes of k_vgemm<KP, MIX>: one k-step of A (all KP / 32 row blocks,
// hi and for split BF16 lo), three buffers, rounded up to whole 256-thread
// passes; at least the epilogue's four 32 x 33 FP32 tiles, which reuse them
// (LDS stores past the allocation are dropped)
template <int KP, int MIX>
constexpr size_t vgemm_lds() {
  return std::max((size_t)3 * (((MIX == MIX_BF16 ? 2 : 1) * (KP / 32) * 64 + 255) / 256 * 256) * 16,
                  (size_t)4 * 32 * 33 * sizeof(float));
}

// A (W or Q) is stored k-step major: fragment (ks, rb, lane) at
// (ks * RB + rb) * 64 + lane, 16 bytes each (FP32: 4 k values for four
// v_mfma_f32_32x32x2_f32; BF16: 8 k values for one v_mfma_f32_32x32x16_bf16),
// so one k-step of every row block is one contiguous RB KB block.  The four
// waves of a workgroup (four 32-column tiles) share it through LDS, in three
// buffers with the global loads two k-steps ahead (one barrier per k-step).
template <int KP, int MIX>
__global__ __launch_bounds__(256) void k_vgemm(VGemmParams p) {
  constexpr int RB = KP / 32;
  constexpr int NPART = (MIX == MIX_BF16) ? 2 : 1;  // hi (, lo)
  constexpr int FR = RB * 64;                        // 16-byte fragments per part and k-step
  constexpr int KS = (MIX == MIX_BF16) ? KP / 16 : KP / 8;
  constexpr int PER = (NPART * FR + 255) / 256;      // fragments each thread stages per k-step
  constexpr int BUF = PER * 256;                     // LDS buffer stride (the tail of the last pass is padding)
  extern __shared__ uint4 sA[];                      // [3][BUF]: [hi | lo | padding]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const VProblem &P = p.prob[blockIdx.y];
  const int ct = blockIdx.x * 4 + wave;
  const bool active = ct * 32 < P.ncols;  // every wave stages and meets the barriers
  f32x16 acc[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[rb][q] = 0.0f;
  const uint4 *const Ahi = (const uint4 *)P.ahi, *const Alo = (const uint4 *)(NPART == 2 ? P.alo : P.ahi);
  // this thread's fragments of a k-step: index tid + 256 i of [hi | lo]; past
  // the end a thread reloads the last fragment and stores it into the padding
  auto src = [&](int idx, size_t o) -> uint4 {
    const int j = idx < NPART * FR ? idx : NPART * FR - 1;
    return j < FR ? Ahi[o + j] : Alo[o + j - FR];
  };
  // this wave's B fragments (its 32 columns); B is streamed once:
  // non-temporal loads keep the re-read A in L2.  A wave past the last tile
  // (ncols >= 1) reads the last tile's fragments and discards them.
  const size_t ctb = (size_t)min(ct, (P.ncols - 1) / 32);
  const uint4 *const Bhi = (const uint4 *)P.bhi + ctb * KS * 64 + lane;
  const uint4 *const Blo = (const uint4 *)(NPART == 2 ? P.blo : P.bhi) + ctb * KS * 64 + lane;
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  auto ntload = [](const uint4 *q) -> uint4 {
    const u32x4 v = __builtin_nontemporal_load((const u32x4 *)q);
    return make_uint4(v.x, v.y, v.z, v.w);
  };
  // Prefetch distance two: k-step ks + 1 is in registers (loaded during
  // k-step ks - 1) while k-step ks is multiplied from LDS; iteration ks
  // stores ks + 1 into LDS buffer (ks + 1) % 3 (last read at ks - 2) and
  // loads ks + 2, so every global load has a whole k-step to arrive.
  // Every k-step issues the same loads (indices past the last k-step are
  // clamped to it; the repeat lands in a buffer no later k-step reads), so
  // the waits count exactly: a k-step waits for the loads of the one before,
  // never for its own.
  uint4 pre[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) pre[i] = src(tid + 256 * i, 0);
#pragma unroll
  for (int i = 0; i < PER; ++i) sA[tid + 256 * i] = pre[i];
#pragma unroll
  for (int i = 0; i < PER; ++i) pre[i] = src(tid + 256 * i, (size_t)min(1, KS - 1) * FR);
  // B of a k-step is loaded one k-step ahead into one of two register sets;
  // the loop is unrolled by two so that the sets alternate by name: no copy
  // of a loaded register at the end of an iteration, whose wait would expose
  // the whole HBM latency of the load in the iteration that issued it.
  uint4 b0h = ntload(Bhi), b0l = (NPART == 2) ? ntload(Blo) : b0h, b1h = b0h, b1l = b0l;
  __syncthreads();
  auto kstep = [&](const int ks, const uint4 bch, const uint4 bcl, uint4 &bnh, uint4 &bnl) {
#pragma unroll
    for (int i = 0; i < PER; ++i) sA[(size_t)((ks + 1) % 3) * BUF + tid + 256 * i] = pre[i];
    const int k1 = min(ks + 1, KS - 1), k2 = min(ks + 2, KS - 1);
#pragma unroll
    for (int i = 0; i < PER; ++i) pre[i] = src(tid + 256 * i, (size_t)k2 * FR);
    bnh = ntload(Bhi + (size_t)k1 * 64);
    if (NPART == 2) bnl = ntload(Blo + (size_t)k1 * 64);
    const uint4 *sh = sA + (size_t)(ks % 3) * BUF;
    if (active) {
      // all of the k-step's A fragments are read from LDS first (one wave per
      // SIMD: the reads' latency overlaps only with each other), then the
      // MFMAs run back to back
      if (MIX == MIX_FP32) {
        const float4 b = make_float4(__uint_as_float(bch.x), __uint_as_float(bch.y), __uint_as_float(bch.z),
                                     __uint_as_float(bch.w));
        uint4 u[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) u[rb] = sh[rb * 64 + lane];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(u[rb].x), b.x, acc[rb], 0, 0, 0);
          acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(u[rb].y), b.y, acc[rb], 0, 0, 0);
          acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(u[rb].z), b.z, acc[rb], 0, 0, 0);
          acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(u[rb].w), b.w, acc[rb], 0, 0, 0);
        }
      } else {
        bf16x8 bh, bl;
        __builtin_memcpy(&bh, &bch, 16);
        __builtin_memcpy(&bl, &bcl, 16);
        uint4 uh[RB], ul[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          uh[rb] = sh[rb * 64 + lane];
          ul[rb] = sh[FR + rb * 64 + lane];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          bf16x8 ah, al;
          __builtin_memcpy(&ah, &uh[rb], 16);
          __builtin_memcpy(&al, &ul[rb], 16);
          acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[rb], 0, 0, 0);
          acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[rb], 0, 0, 0);
          acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[rb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  };
  int ks = 0;
  for (; ks + 1 < KS; ks += 2) {
    kstep(ks, b0h, b0l, b1h, b1l);
    kstep(ks + 1, b1h, b1l, b0h, b0l);
  }
  if (ks < KS) kstep(ks, b0h, b0l, b1h, b1l);
  if (!active) return;
  // C(row, col) = acc[rb][q], row = 32 rb + 8 (q / 4) + 4 (lane / 32) + q % 4,
  // col = 32 ct + lane % 32.  Column pair m = col >> 1 (Re, Im of one phi:
  // m = (pair * NE + Euler) * nphi + phi).  Each 32 x 32 block of C goes
  // through this wave's LDS tile T[col][row] (the A buffers are free now),
  // so that the FP64 vectors of the dot epilogue are read, and G is written,
  // in contiguous runs of 32 rows per column pair.
  const int hh = p.h * p.h;
  float *T = (float *)sA + (size_t)wave * 32 * 33;
  // this lane's (pair, row) share of a block: pairs mp = lane / 32 + 2 i, row lane % 32
  const int rl = lane & 31;
  if (P.mode == 0) {
    double sr[8], si[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) sr[i] = si[i] = 0.0;
    // the FP64 vector of each of this lane's column pairs (pairs past ncols
    // read the last pair's, with weight 0): no load sits in a branch, so the
    // loads of a block are issued together instead of one wait per load
    const double2 *dcol[8];
    double wcol[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const size_t m0 = (size_t)ct * 16 + (lane >> 5) + 2 * i, mlast = (size_t)(P.ncols / 2 - 1);
      const size_t m = m0 < mlast ? m0 : mlast;
      dcol[i] = P.dot + (((m / p.nphi) * 2 + P.species) * p.nphi + (m % p.nphi)) * hh;
      wcol[i] = (2 * m0 < (size_t)P.ncols) ? 1.0 : 0.0;
    }
    // rows past h^2 (the K padding of the last block) read row h^2 - 1 with
    // weight 0; the vectors of block rb + DP are loaded while block rb is
    // summed (the accumulators' registers are free after the k-loop)
    constexpr int DP = RB < 4 ? RB : 4;
    double2 dv[DP][8];
    // (the scheduling barrier keeps the compiler from sinking the loads to
    // their uses, which would wait on each load again)
    auto dload = [&](double2 *d, int rb) {
      const int row = 32 * rb + rl, rowc = row < hh ? row : hh - 1;
#pragma unroll
      for (int i = 0; i < 8; ++i) d[i] = dcol[i][rowc];
      __builtin_amdgcn_sched_barrier(0);
    };
#pragma unroll
    for (int rb = 0; rb < DP; ++rb) dload(dv[rb], rb);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
      for (int q = 0; q < 16; ++q) T[(lane & 31) * 33 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)] = acc[rb][q];
      const double wr = (32 * rb + rl < hh) ? 1.0 : 0.0;
      const double2 *d = dv[rb % DP];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int mp = (lane >> 5) + 2 * i;
        const double w = wr * wcol[i];
        const double g0 = w * (double)T[(2 * mp) * 33 + rl], g1 = w * (double)T[(2 * mp + 1) * 33 + rl];
        sr[i] += d[i].x * g0 - d[i].y * g1;  // d (G_re + i G_im)
        si[i] += d[i].y * g0 + d[i].x * g1;
      }
      if (rb + DP < RB) dload(dv[rb % DP], rb + DP);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int off = 16; off > 0; off >>= 1) {
        sr[i] += __shfl_xor(sr[i], off);
        si[i] += __shfl_xor(si[i], off);
      }
    if (rl == 0)
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const size_t m = (size_t)ct * 16 + (lane >> 5) + 2 * i;
        if (2 * m < (size_t)P.ncols) {
          const size_t ob = ((m / p.nphi) * 2 + P.species) * p.nphi + (m % p.nphi);
          p.out[ob * FAC_OUT + P.slot] = make_double2(sr[i], si[i]);
        }
      }
  } else {
    float2 *g = (float2 *)P.g;  // G[m][row] = (Re, Im)
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
      for (int q = 0; q < 16; ++q) T[(lane & 31) * 33 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)] = acc[rb][q];
      const int row = 32 * rb + rl;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int mp = (lane >> 5) + 2 * i;
        const size_t m = (size_t)ct * 16 + mp;
        if (2 * m < (size_t)P.ncols && row < hh)
          g[m * hh + row] = make_float2(T[(2 * mp) * 33 + rl], T[(2 * mp + 1) * 33 + rl]);
      }
    }
  }
}

template <int KP>
static void launch_vgemm_kp(int mix, dim3 grid, hipStream_t s, const VGemmParams &vp) {
  constexpr size_t l32 = vgemm_lds<KP, MIX_FP32>(), l16 = vgemm_lds<KP, MIX_BF16>();
  if (mix == MIX_FP32) hipLaunchKernelGGL((k_vgemm<KP, MIX_FP32>), grid, dim3(256), l32, s, vp);
  else hipLaunchKernelGGL((k_vgemm<KP, MIX_BF16>), grid, dim3(256), l16, s, vp);
}

// Dynamic LDS limits of both precisions (set once per device, before any capture).
template <int KP>
static hipError_t vgemm_attr_kp() {
  hipError_t e = hipFuncSetAttribute((const void *)k_vgemm<KP, MIX_FP32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)vgemm_lds<KP, MIX_FP32>());
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_vgemm<KP, MIX_BF16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)vgemm_lds<KP, MIX_BF16>());
  return e;
}

static hipError_t vgemm_attr() {
  const hipError_t e[] = {vgemm_attr_kp<32>(), vgemm_attr_kp<64>(), vgemm_attr_kp<128>(), vgemm_attr_kp<160>(),
                          vgemm_attr_kp<224>(), vgemm_attr_kp<256>(), vgemm_attr_kp<352>(), vgemm_attr_kp<416>(),
                          vgemm_attr_kp<512>()};
  for (hipError_t x : e)
    if (x != hipSuccess) return x;
  return hipSuccess;
}

// KP = h^2 rounded up to 32, for the instantiated species sizes
static bool launch_vgemm(int h, int mix, dim3 grid, hipStream_t s, const VGemmParams &vp) {
  switch (h) {
    case 2: case 4: launch_vgemm_kp<32>(mix, grid, s, vp); return true;
    case 6: launch_vgemm_kp<64>(mix, grid, s, vp); return true;
    case 8: launch_vgemm_kp<64>(mix, grid, s, vp); return true;
    case 10: launch_vgemm_kp<128>(mix, grid, s, vp); return true;
    case 12: launch_vgemm_kp<160>(mix, grid, s, vp); return true;
    case 14: launch_vgemm_kp<224>(mix, grid, s, vp); return true;
    case 16: launch_vgemm_kp<256>(mix, grid, s, vp); return true;
    case 18: launch_vgemm_kp<352>(mix, grid, s, vp); return true;
    case 20: launch_vgemm_kp<416>(mix, grid, s, vp); return true;
    case 22: launch_vgemm_kp<512>(mix, grid, s, vp); return true;
    default: return false;
  }
}

// Occupancy variants of the FP64 factored form that the autotuner times
// (species sizes h >= 8; smaller spaces take the defaults): Pfaffian part at
// three, four or five waves per SIMD, the rest at its default or, for
// h <= 16, four (a tighter register cap).  0 = the default.
struct FacTune {
  int pf_waves = 0, rest_waves = 0, chunk_div = 1;
};

// route: 0 the 2h x 2h Pfaffian, 1 the Schur route for every state, 2 both
// kernels (pp.pfs_ok names each pair's); a pf_waves choice takes the 2h x 2h one
template <int H>
static void launch_phi(int mix, const FacTune &t, int route, unsigned grid, hipStream_t s, const PhiParams &pp) {
  const dim3 g(grid), b(64);
  if (route >= 1 && t.pf_waves == 0) hipLaunchKernelGGL((k_fac_phi<H, MIX_FP64, PART_PFS>), g, b, 0, s, pp);
  if (route == 1 && t.pf_waves == 0) {
  } else if (H >= 8 && mix == MIX_FP64 && t.pf_waves == 3) hipLaunchKernelGGL((k_fac_phi<H, MIX_FP64, PART_PF, 3>), g, b, 0, s, pp);
  else if (H >= 8 && mix == MIX_FP64 && t.pf_waves == 4) hipLaunchKernelGGL((k_fac_phi<H, MIX_FP64, PART_PF, 4>), g, b, 0, s, pp);
  else if (H >= 8 && mix == MIX_FP64 && t.pf_waves == 5) hipLaunchKernelGGL((k_fac_phi<H, MIX_FP64, PART_PF, 5>), g, b, 0, s, pp);
  else hipLaunchKernelGGL((k_fac_phi<H, MIX_FP64, PART_PF>), g, b, 0, s, pp);
  if (mix == MIX_FP32) hipLaunchKernelGGL((k_fac_phi<H, MIX_FP32, PART_REST>), g, b, 0, s, pp);
  else if (mix == MIX_BF16) hipLaunchKernelGGL((k_fac_phi<H, MIX_BF16, PART_REST>), g, b, 0, s, pp);
  else if (H >= 8 && H <= 16 && t.rest_waves == 4)
    hipLaunchKernelGGL((k_fac_phi<H, MIX_FP64, PART_REST, 4>), g, b, 0, s, pp);
  else hipLaunchKernelGGL((k_fac_phi<H, MIX_FP64, PART_REST>), g, b, 0, s, pp);
}

// Scratch (register spill) bytes per lane of the FP64 variant kernels a
// tuning choice launches; the autotuner only takes spill-free variants.
template <int H>
static size_t phi_scratch(const FacTune &t, bool schur) {
  const void *pf = (const void *)k_fac_phi<H, MIX_FP64, PART_PF>;
  if (schur && t.pf_waves == 0) pf = (const void *)k_fac_phi<H, MIX_FP64, PART_PFS>;
  if (H >= 8 && t.pf_waves == 3) pf = (const void *)k_fac_phi<H, MIX_FP64, PART_PF, 3>;
  if (H >= 8 && t.pf_waves == 4) pf = (const void *)k_fac_phi<H, MIX_FP64, PART_PF, 4>;
  if (H >= 8 && t.pf_waves == 5) pf = (const void *)k_fac_phi<H, MIX_FP64, PART_PF, 5>;
  const void *rest = (const void *)k_fac_phi<H, MIX_FP64, PART_REST>;
  if (H >= 8 && H <= 16 && t.rest_waves == 4) rest = (const void *)k_fac_phi<H, MIX_FP64, PART_REST, 4>;
  hipFuncAttributes a{}, b{};
  if (hipFuncGetAttributes(&a, pf) != hipSuccess || hipFuncGetAttributes(&b, rest) != hipSuccess) return (size_t)-1;
  return a.localSizeBytes + b.localSizeBytes;
}

static size_t phi_scratch_h(int h, const FacTune &t, bool schur) {
  switch (h) {
    case 2: return phi_scratch<2>(t, schur);
    case 4: return phi_scratch<4>(t, schur);
    case 6: return phi_scratch<6>(t, schur);
    case 8: return phi_scratch<8>(t, schur);
    case 10: return phi_scratch<10>(t, schur);
    case 12: return phi_scratch<12>(t, schur);
    case 14: return phi_scratch<14>(t, schur);
    case 16: return phi_scratch<16>(t, schur);
    case 18: return phi_scratch<18>(t, schur);
    case 20: return phi_scratch<20>(t, schur);
    case 22: return phi_scratch<22>(t, schur);
    default: return (size_t)-1;
  }
}

// Species sizes h = n / 2 with an instantiated k_fac_phi (every h is even:
// an orbit holds 2j + 1 states).
static bool launch_phi_h(int h, int mix, const FacTune &t, int route, unsigned grid, hipStream_t s,
                         const PhiParams &pp) {
  switch (h) {
    case 2: launch_phi<2>(mix, t, route, grid, s, pp); return true;
    case 4: launch_phi<4>(mix, t, route, grid, s, pp); return true;
    case 6: launch_phi<6>(mix, t, route, grid, s, pp); return true;
    case 8: launch_phi<8>(mix, t, route, grid, s, pp); return true;
    case 10: launch_phi<10>(mix, t, route, grid, s, pp); return true;
    case 12: launch_phi<12>(mix, t, route, grid, s, pp); return true;
    case 14: launch_phi<14>(mix, t, route, grid, s, pp); return true;
    case 16: launch_phi<16>(mix, t, route, grid, s, pp); return true;
    case 18: launch_phi<18>(mix, t, route, grid, s, pp); return true;
    case 20: launch_phi<20>(mix, t, route, grid, s, pp); return true;
    case 22: launch_phi<22>(mix, t, route, grid, s, pp); return true;
    default: return false;
  }
}

template <int NP>
static void launch_combine(bool gmix, dim3 grid, size_t lds, hipStream_t s, const CombParams &cp) {
  if (gmix) hipLaunchKernelGGL((k_fac_combine<NP, true>), grid, dim3(NTC), lds, s, cp);
  else hipLaunchKernelGGL((k_fac_combine<NP, false>), grid, dim3(NTC), lds, s, cp);
}

// Dynamic LDS limit of k_fac_combine<np, *>, set before any capture.
template <int NP>
static bool combine_attr_np(size_t lds) {
  return hipFuncSetAttribute((const void *)k_fac_combine<NP, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds) == hipSuccess &&
         hipFuncSetAttribute((const void *)k_fac_combine<NP, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds) == hipSuccess;
}
static bool combine_attr(int np, size_t lds) {
  switch (np) {
    case 1: return combine_attr_np<1>(lds);
    case 2: return combine_attr_np<2>(lds);
    case 3: return combine_attr_np<3>(lds);
    case 4: return combine_attr_np<4>(lds);
    case 5: return combine_attr_np<5>(lds);
    case 6: return combine_attr_np<6>(lds);
    case 7: return combine_attr_np<7>(lds);
    case 8: return combine_attr_np<8>(lds);
    case 9: return combine_attr_np<9>(lds);
    case 10: return combine_attr_np<10>(lds);
    case 11: return combine_attr_np<11>(lds);
    case 12: return combine_attr_np<12>(lds);
    case 13: return combine_attr_np<13>(lds);
    case 14: return combine_attr_np<14>(lds);
    case 15: return combine_attr_np<15>(lds);
    case 16: return combine_attr_np<16>(lds);
    default: return false;
  }
}

static bool launch_combine_np(int np, bool gmix, dim3 grid, size_t lds, hipStream_t s, const CombParams &cp) {
  switch (np) {
    case 1: launch_combine<1>(gmix, grid, lds, s, cp); return true;
    case 2: launch_combine<2>(gmix, grid, lds, s, cp); return true;
    case 3: launch_combine<3>(gmix, grid, lds, s, cp); return true;
    case 4: launch_combine<4>(gmix, grid, lds, s, cp); return true;
    case 5: launch_combine<5>(gmix, grid, lds, s, cp); return true;
    case 6: launch_combine<6>(gmix, grid, lds, s, cp); return true;
    case 7: launch_combine<7>(gmix, grid, lds, s, cp); return true;
    case 8: launch_combine<8>(gmix, grid, lds, s, cp); return true;
    case 9: launch_combine<9>(gmix, grid, lds, s, cp); return true;
    case 10: launch_combine<10>(gmix, grid, lds, s, cp); return true;
    case 11: launch_combine<11>(gmix, grid, lds, s, cp); return true;
    case 12: launch_combine<12>(gmix, grid, lds, s, cp); return true;
    case 13: launch_combine<13>(gmix, grid, lds, s, cp); return true;
    case 14: launch_combine<14>(gmix, grid, lds, s, cp); return true;
    case 15: launch_combine<15>(gmix, grid, lds, s, cp); return true;
    case 16: launch_combine<16>(gmix, grid, lds, s, cp); return true;
    default: return false;
  }
}

// ------------------------------------------------------------------ projection
struct ProjParams {
  int nalp, nbet, ngam, NE, jmax, fold, npj;
  int nobs;                  // observables per Euler point in zres (and per pair in kern)
  const double *alp, *walp, *wbet, *gam, *wgam;
  const double *ptab;        // [nbet][2][npj]
  const double2 *zres;       // [npairs][NE][NOBS]
  const int *status;         // [npairs][NE]
  double2 *kern;             // [npairs][NOBS][jmax+1][2jmax+1][2jmax+1]
  int *pair_status;          // [npairs]
};

// The phases e^{i K alpha_a} and e^{i K gamma_g} (K = -jmax..jmax) are tabled
// in LDS (dynamic, (2 jmax + 1)(nalp + ngam) complex), so a term of the sum is
// one complex product instead of a sine and cosine.
__global__ __launch_bounds__(256) void k_project(ProjParams p) {
  extern __shared__ double2 ktab[];
  const int ip = blockIdx.x, io = blockIdx.y, tid = threadIdx.x;
  const int J2m = 2 * p.jmax + 1;
  double2 *ea = ktab, *eg = ktab + (size_t)J2m * p.nalp;
  for (int i = tid; i < J2m * p.nalp; i += blockDim.x) ea[i] = cexpi((i / p.nalp - p.jmax) * p.alp[i % p.nalp]);
  for (int i = tid; i < J2m * p.ngam; i += blockDim.x) eg[i] = cexpi((i / p.ngam - p.jmax) * p.gam[i % p.ngam]);
  __syncthreads();
  const double2 *z = p.zres + (size_t)ip * p.NE * p.nobs;
  double2 *kout = p.kern + ((size_t)ip * p.nobs + io) * (p.jmax + 1) * J2m * J2m;
  const bool onept = (p.nalp * p.ngam == 1);
  const int ntot = (p.jmax + 1) * J2m * J2m;
  for (int o = tid; o < ntot; o += blockDim.x) {
    const int J = o / (J2m * J2m), K1 = (o / J2m) % J2m - p.jmax, K2 = o % J2m - p.jmax;
    if (K1 < -J || K1 > J || K2 < -J || K2 > J) {
      kout[o] = cz();
      continue;
    }
    // compact d-table index: sum_{J'<J} (2J'+1)^2 = J (4J^2 - 1) / 3
    const int e = J * (4 * J * J - 1) / 3 + (K1 + J) * (2 * J + 1) + (K2 + J);
    double2 acc = cz();
    if (!(p.fold == PNAMP_FOLD_LITERAL && (J & 1))) {
      const double wJ = onept ? (2.0 * J + 1.0) / 2.0
                              : (p.fold == PNAMP_FOLD_LITERAL ? 2.0 : 1.0) * (2.0 * J + 1.0) / (8.0 * M_PI * M_PI);
      const double d0 = (K1 == 0 ? 2.0 : 1.0) * (K2 == 0 ? 2.0 : 1.0);
      for (int ib = 0; ib < p.nbet; ++ib) {
        const double dp = p.ptab[((size_t)ib * 2 + 0) * p.npj + e];
        const double dm = p.ptab[((size_t)ib * 2 + 1) * p.npj + e];
        for (int ia = 0; ia < p.nalp; ++ia)
          for (int ig = 0; ig < p.ngam; ++ig) {
            const double w = p.walp[ia] * p.wbet[ib] * p.wgam[ig];
            const int i1 = (ib * p.nalp + ia) * p.ngam + ig;
            const double2 zap = ea[(K1 + p.jmax) * p.nalp + ia], zgp = eg[(K2 + p.jmax) * p.ngam + ig];
            if (p.fold == PNAMP_FOLD_PLAIN) {
              const double2 D = cscale(dp * w * wJ, cmul(zap, zgp));
              cfma(acc, D, z[(size_t)i1 * p.nobs + io]);
            } else {
              const int i2 = (ib * p.nalp + ia) * p.ngam + (p.ngam - 1 - ig);
              const int i3 = (ib * p.nalp + (p.nalp - 1 - ia)) * p.ngam + ig;
              const int i4 = (ib * p.nalp + (p.nalp - 1 - ia)) * p.ngam + (p.ngam - 1 - ig);
              const double2 zam = cconj(zap);
              const double2 za1 = cscale(dp, cmul(zap, zgp)), za2 = cscale(dp, cmul(zam, zgp));
              const double2 za3 = cscale(dm, cmul(zap, zgp)), za4 = cscale(dm, cmul(zam, zgp));
              double2 s = cmul(za1, z[(size_t)i1 * p.nobs + io]);
              s = cadd(s, cmul(za2, cconj(z[(size_t)i2 * p.nobs + io])));
              s = cadd(s, cmul(za3, cconj(z[(size_t)i3 * p.nobs + io])));
              s = cadd(s, cmul(za4, z[(size_t)i4 * p.nobs + io]));
              acc = cadd(acc, cscale(wJ * w / d0, s));
            }
          }
      }
    }
    kout[o] = acc;
  }
  if (io == 0) {
    __shared__ int st[256];
    int s = 0;
    for (int i = tid; i < p.NE; i += blockDim.x) {
      const int v = p.status[(size_t)ip * p.NE + i];
      if (v && !s) s = v;
    }
    st[tid] = s;
    __syncthreads();
    if (tid == 0) {
      int r = 0;
      for (int i = 0; i < (int)blockDim.x; ++i)
        if (st[i]) { r = st[i]; break; }
      p.pair_status[ip] = r;
    }
  }
}

// Plain fold on the product mesh, separable: per beta node the gamma sum
// (per K2), then the alpha sum (per K1, K2), then the d^J_{K1K2}(beta) term,
// (2J+1) / (8 pi^2) sum_abg w_a w_b w_g e^{iK1 a} d^J_{K1K2}(b) e^{iK2 g} ZRES
// in nbet (ngam + nalp (2jmax+1)) (2jmax+1) + nbet npj complex products
// instead of nalp nbet ngam npj.  Up to PROJ_MAXO outputs per thread.
constexpr int PROJ_MAXO = 16;
__global__ __launch_bounds__(256) void k_project_sep(ProjParams p) {
  extern __shared__ double2 ktab[];
  const int ip = blockIdx.x, io = blockIdx.y, tid = threadIdx.x;
  const int J2m = 2 * p.jmax + 1, na = p.nalp, ng = p.ngam;
  double2 *ea = ktab, *eg = ea + (size_t)J2m * na, *zb = eg + (size_t)J2m * ng, *zg = zb + (size_t)na * ng,
          *zag = zg + (size_t)na * J2m;
  for (int i = tid; i < J2m * na; i += blockDim.x) ea[i] = cexpi((i / na - p.jmax) * p.alp[i % na]);
  for (int i = tid; i < J2m * ng; i += blockDim.x) eg[i] = cexpi((i / ng - p.jmax) * p.gam[i % ng]);
  const double2 *z = p.zres + (size_t)ip * p.NE * p.nobs;
  double2 *kout = p.kern + ((size_t)ip * p.nobs + io) * (p.jmax + 1) * J2m * J2m;
  const int ntot = (p.jmax + 1) * J2m * J2m;
  double2 acc[PROJ_MAXO];
#pragma unroll
  for (int m = 0; m < PROJ_MAXO; ++m) acc[m] = cz();
  for (int ib = 0; ib < p.nbet; ++ib) {
    __syncthreads();
    for (int i = tid; i < na * ng; i += blockDim.x) zb[i] = z[((size_t)ib * na * ng + i) * p.nobs + io];
    __syncthreads();
    for (int i = tid; i < na * J2m; i += blockDim.x) {
      const int a = i / J2m, k2 = i % J2m;
      double2 sacc = cz();
      for (int g = 0; g < ng; ++g) cfma(sacc, cscale(p.wgam[g], eg[k2 * ng + g]), zb[a * ng + g]);
      zg[i] = sacc;
    }
    __syncthreads();
    for (int i = tid; i < J2m * J2m; i += blockDim.x) {
      const int k1 = i / J2m, k2 = i % J2m;
      double2 sacc = cz();
      for (int a = 0; a < na; ++a) cfma(sacc, cscale(p.walp[a], ea[k1 * na + a]), zg[a * J2m + k2]);
      zag[i] = sacc;
    }
    __syncthreads();
    const double wb = p.wbet[ib];
#pragma unroll
    for (int m = 0; m < PROJ_MAXO; ++m) {
      const int o = tid + m * (int)blockDim.x;
      if (o >= ntot) break;
      const int J = o / (J2m * J2m), K1 = (o / J2m) % J2m - p.jmax, K2 = o % J2m - p.jmax;
      if (K1 < -J || K1 > J || K2 < -J || K2 > J) continue;
      const int e = J * (4 * J * J - 1) / 3 + (K1 + J) * (2 * J + 1) + (K2 + J);
      const double dp = p.ptab[((size_t)ib * 2 + 0) * p.npj + e];
      acc[m] = cadd(acc[m], cscale(wb * dp, zag[(K1 + p.jmax) * J2m + (K2 + p.jmax)]));
    }
  }
  const bool onept = (na * ng == 1);
#pragma unroll
  for (int m = 0; m < PROJ_MAXO; ++m) {
    const int o = tid + m * (int)blockDim.x;
    if (o >= ntot) break;
    const int J = o / (J2m * J2m);
    const double wJ = onept ? (2.0 * J + 1.0) / 2.0 : (2.0 * J + 1.0) / (8.0 * M_PI * M_PI);
    kout[o] = cscale(wJ, acc[m]);
  }
  if (io == 0) {
    __shared__ int st[256];
    int sv = 0;
    for (int i = tid; i < p.NE; i += blockDim.x) {
      const int v = p.status[(size_t)ip * p.NE + i];
      if (v && !sv) sv = v;
    }
    st[tid] = sv;
    __syncthreads();
    if (tid == 0) {
      int r = 0;
      for (int i = 0; i < (int)blockDim.x; ++i)
        if (st[i]) { r = st[i]; break; }
      p.pair_status[ip] = r;
    }
  }
}

// ------------------------------------------------------------------ Hill-Wheeler
// Cyclic Jacobi on a real symmetric dim x dim matrix A (global memory,
// column-major) by the calling workgroup, round-robin pairing so that dim/2
// rotations run in parallel per step; eigenvectors accumulate in Vv.  Ends
// after a sweep whose off-diagonal norm is below 1e-30 of the total, or after
// 60 sweeps; returns the sweep count.  Scratch in LDS: cs, sn, pp, qq of 512
// entries (dim <= 1022), red of blockDim.x entries.
struct JacobiLds {
  double cs[512], sn[512];
  int pp[512], qq[512];
  double red[1024];  // one partial sum per thread (up to 1024 threads)
  int done;
};

__device__ int block_jacobi(double *A, double *Vv, int dim, JacobiLds &L) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int m = dim + (dim & 1);  // even number of slots, slot dim is a dummy
  for (int e = tid; e < dim * dim; e += nt) Vv[e] = ((e % dim) == (e / dim)) ? 1.0 : 0.0;
  __syncthreads();
  int sweep = 0;
  for (sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < dim * dim; e += nt) {
      const double x = A[e] * A[e];
      tot += x;
      if ((e % dim) != (e / dim)) off += x;
    }
    L.red[tid] = off;
    __syncthreads();
    if (tid == 0) {
      double o = 0.0;
      for (int i = 0; i < nt; ++i) o += L.red[i];
      L.red[0] = o;
    }
    __syncthreads();
    const double offs = L.red[0];
    __syncthreads();
    L.red[tid] = tot;
    __syncthreads();
    if (tid == 0) {
      double o = 0.0;
      for (int i = 0; i < nt; ++i) o += L.red[i];
      L.done = (offs <= 1e-30 * o || offs == 0.0);
    }
    __syncthreads();
    if (L.done) break;
    for (int step = 0; step < m - 1; ++step) {
      // circle method: slot 0 fixed, slots 1..m-1 on a circle of odd length
      for (int k = tid; k < m / 2; k += nt) {
        int a, b;
        if (k == 0) {
          a = 0;
          b = 1 + step % (m - 1);
        } else {
          a = 1 + (step + k) % (m - 1);
          b = 1 + (step - k + (m - 1)) % (m - 1);
        }
        const int P = a < b ? a : b, Q = a < b ? b : a;
        double c = 1.0, s = 0.0;
        if (Q < dim) {
          const double apq = A[P + Q * dim];
          if (fabs(apq) > 1e-300) {
            const double th = (A[Q + Q * dim] - A[P + P * dim]) / (2.0 * apq);
            const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
          }
        }
        L.pp[k] = P; L.qq[k] = Q; L.cs[k] = c; L.sn[k] = s;
      }
      __syncthreads();
      // A = J^T A J: one thread per 2 x 2 block (rows of pair a, columns of
      // pair b).  The column and the row rotation of a block touch only its
      // four entries, so every block updates in place on its own (columns
      // first, then rows, as two separate passes over the matrix would);
      // pairs with the dummy slot have c = 1, s = 0 and no dummy entries.
      const int hp = m / 2;
      for (int e = tid; e < hp * hp; e += nt) {
        const int a = e % hp, b = e / hp;
        const int Pa = L.pp[a], Qa = L.qq[a], Pb = L.pp[b], Qb = L.qq[b];
        const bool qa = Qa < dim, qb = Qb < dim;
        const double ca = L.cs[a], sa = L.sn[a], cb = L.cs[b], sb = L.sn[b];
        const double xpp = A[Pa + Pb * dim];
        const double xpq = qb ? A[Pa + Qb * dim] : 0.0;
        const double xqp = qa ? A[Qa + Pb * dim] : 0.0;
        const double xqq = (qa && qb) ? A[Qa + Qb * dim] : 0.0;
        const double ypp = cb * xpp - sb * xpq, ypq = sb * xpp + cb * xpq;  // columns: A J
        const double yqp = cb * xqp - sb * xqq, yqq = sb * xqp + cb * xqq;
        A[Pa + Pb * dim] = ca * ypp - sa * yqp;  // rows: J^T (A J)
        if (qb) A[Pa + Qb * dim] = ca * ypq - sa * yqq;
        if (qa) A[Qa + Pb * dim] = sa * ypp + ca * yqp;
        if (qa && qb) A[Qa + Qb * dim] = sa * ypq + ca * yqq;
      }
      // V = V J (columns P, Q of every pair; consecutive threads, consecutive rows)
      for (int e = tid; e < hp * dim; e += nt) {
        const int k = e / dim, r = e % dim, P = L.pp[k], Q = L.qq[k];
        if (Q >= dim) continue;
        const double c = L.cs[k], s = L.sn[k];
        const double vp = Vv[r + P * dim], vq = Vv[r + Q * dim];
        Vv[r + P * dim] = c * vp - s * vq;
        Vv[r + Q * dim] = s * vp + c * vq;
      }
      __syncthreads();
    }
  }
  return sweep;
}

// (one workgroup of 1024 threads: the matrix streams through L2 every step,
// so the workgroup's memory parallelism bounds it)
__global__ __launch_bounds__(1024) void k_jacobi(double *A, double *Vv, int dim, int *nsweeps) {
  __shared__ JacobiLds L;
  const int sw = block_jacobi(A, Vv, dim, L);
  if (threadIdx.x == 0) *nsweeps = sw;
}

// Natural-state scan of PreSolve_HWG (Codes/GCM/GCM_HWG.f90:2-100): block b
// solves HNDIAG restricted to the first nos = b + 1 natural states.  The
// columns of R (dim x M, R(:, k) = D(:, k) / sqrt(EN_k), EN descending) are
// nested, so the reduced Hamiltonian of nos states is the leading nos x nos
// block of G = R^T H R (M x M).  Per level k < min(kmax, nos): the energy
// E_k (ascending), the amplitudes f = R(:, :nos) g_k (FF of HNDIAG) and the
// variance q2 = sum_i [sum_j (H_ij - E_k N_ij) f_j]^2 of mean_variance
// (:205-254).  work holds nos^2 + nos^2 doubles per block at offset
// 2 sum_{i <= b} i^2; E and var are [M][kmax].
struct NosParams {
  int dim, M, kmax;
  const double *N, *H, *R, *G;
  double *work, *E, *var;
};

__global__ __launch_bounds__(256) void k_nos_scan(NosParams p) {
  __shared__ JacobiLds L;
  __shared__ int ord[1024];
  __shared__ double f[1024], z[1024];
  const int b = blockIdx.x, nos = b + 1, tid = threadIdx.x, nt = blockDim.x, dim = p.dim;
  const size_t woff = 2 * ((size_t)b * (b + 1) * (2 * b + 1) / 6);
  double *A = p.work + woff, *Vv = A + (size_t)nos * nos;
  for (int e = tid; e < nos * nos; e += nt) A[e] = p.G[(e % nos) + (size_t)(e / nos) * p.M];
  __syncthreads();
  block_jacobi(A, Vv, nos, L);
  // ascending order of the eigenvalues (ties by index)
  for (int i = tid; i < nos; i += nt) {
    const double ei = A[i + (size_t)i * nos];
    int r = 0;
    for (int j = 0; j < nos; ++j) {
      const double ej = A[j + (size_t)j * nos];
      r += (ej < ei || (ej == ei && j < i)) ? 1 : 0;
    }
    ord[r] = i;
  }
  __syncthreads();
  const int kk = min(p.kmax, nos);
  for (int k = 0; k < p.kmax; ++k) {
    double *Eo = p.E + (size_t)b * p.kmax + k, *vo = p.var + (size_t)b * p.kmax + k;
    if (k >= kk) {
      if (tid == 0) { *Eo = 0.0; *vo = 0.0; }
      continue;
    }
    const int col = ord[k];
    const double ek = A[col + (size_t)col * nos];
    for (int i = tid; i < dim; i += nt) {
      double s = 0.0;
      for (int l = 0; l < nos; ++l) s = fma(p.R[i + (size_t)l * dim], Vv[l + (size_t)col * nos], s);
      f[i] = s;
    }
    __syncthreads();
    double q = 0.0;
    for (int i = tid; i < dim; i += nt) {
      double s = 0.0;
      for (int j = 0; j < dim; ++j) s = fma(p.H[i + (size_t)j * dim] - ek * p.N[i + (size_t)j * dim], f[j], s);
      z[i] = s;
      q = fma(s, s, q);
    }
    L.red[tid] = q;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int i = 0; i < nt; ++i) t += L.red[i];
      *Eo = ek;
      *vo = t;
    }
    __syncthreads();
  }
}

// W = R^T H R for R (dim x M) and symmetric H (dim x dim): one thread per (j, k)
__global__ void k_congruence(const double *H, const double *R, double *W, int dim, int M) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * M) return;
  const int j = e % M, k = e / M;
  double s = 0.0;
  for (int l = 0; l < dim; ++l) {
    double hr = 0.0;
    for (int i = 0; i < dim; ++i) hr = fma(H[l + i * dim], R[i + k * dim], hr);
    s = fma(R[l + j * dim], hr, s);
  }
  W[j + k * M] = s;
}

__global__ void k_scale_cols(const double *D, const double *w, double *R, int dim, int M) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= dim * M) return;
  const int i = e % dim, k = e / dim;
  R[i + k * dim] = D[i + k * dim] / sqrt(w[k]);
}

__global__ void k_matmul(const double *A, const double *B, double *C, int m, int kk, int nc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * nc) return;
  const int i = e % m, j = e / m;
  double s = 0.0;
  for (int l = 0; l < kk; ++l) s = fma(A[i + l * m], B[l + j * kk], s);
  C[i + j * m] = s;
}

}  // namespace

// Device memory of pnamp_kernels_batched_noalloc: every intermediate of that
// entry point (pair lists, ZRES, status flags, projected kernels) lives here.
// Each GPU that loads the library holds its own copy until the process ends;
// host code reaches it only through the per-device address from
// hipGetSymbolAddress (noalloc_scratch_ptr), never through the name.
static __device__ __align__(256) unsigned char scratch[PNAMP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Address of this device's scratch array; looked up once per device (with
// that device current).  nullptr and a message on failure.
void *noalloc_scratch_ptr(int device, std::string &err) {
  if (device < 0 || device >= MAX_DEVICES) {
    err = "device index beyond the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}
}  // namespace

// ====================================================================== host
struct pnamp_ctx {
  int device = 0;
  void *scratch = nullptr;  // this device's scratch array (pnamp_kernels_batched_noalloc)
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  float ms_points = 0.f, ms_total = 0.f;
  std::string err;
  // basis
  int n = 0, h = 0, npack = 0;
  std::vector<int> twom_h;
  // mesh
  int nalp = 0, nbet = 0, ngam = 0, nphi = 0, jmax = 0, fold = 0, NE = 0, npj = 0;
  double cscale = 1.0, vac_scale = 1.0, snpfaf = 1.0, e0 = 0.0;
  int nme = 0;
  int nstates = 0;
  int pair_cap = 0;
  // device buffers
  double *d_alp = nullptr, *d_walp = nullptr, *d_bet = nullptr, *d_wbet = nullptr, *d_gam = nullptr,
         *d_wgam = nullptr;
  double *d_rtab = nullptr, *d_ptab = nullptr;
  int *d_pk_a = nullptr, *d_pk_b = nullptr, *d_blk_off = nullptr, *d_blk_sz = nullptr,
      *d_blk_base = nullptr, *d_twoj = nullptr, *d_twom = nullptr;
  // plain form: the regrouped energy list (LDS offsets, matrix elements) and
  // the work buffer of k_plain_* (units of the current chunk)
  uint4 *d_plofs = nullptr;
  double *d_plme = nullptr;
  int pl_n = 0, pl_bytes = 0;
  void *d_pwork = nullptr;
  size_t pwork_bytes = 0;
  double *d_t1b = nullptr;
  double *d_jtab = nullptr;
  double *d_U = nullptr, *d_V = nullptr, *d_Z = nullptr, *d_det = nullptr;
  int *d_state_status = nullptr;
  int *d_q0 = nullptr, *d_q1 = nullptr, *d_pz = nullptr, *d_pn = nullptr;
  double2 *d_zres = nullptr, *d_kern = nullptr;
  int *d_status = nullptr, *d_pair_status = nullptr;
  // factored form
  // E2 observables (pnamp_set_q2): Q_{2mu}, orbit pairs, 3j table and the
  // per-call outputs (ZRES of the ten E2 observables and the orbit traces,
  // their projections, reduced elements, J = 0 one-body density)
  std::vector<int> twoj_h, l_h, nq_h;
  double *d_q2 = nullptr, *d_t3j = nullptr, *d_orbscale = nullptr;
  int4 *d_orbp = nullptr;
  int nob2 = 0, norb = 0;
  double2 *d_ze2 = nullptr, *d_zrho = nullptr, *d_iproj = nullptr, *d_qred = nullptr, *d_rho1b = nullptr;
  int e2_cap = 0;
  int form = PNAMP_FORM_PLAIN;
  bool separable = false;
  int nme_s[3] = {0, 0, 0};
  uint4 *d_lofs[2] = {nullptr, nullptr};            // species lists of k_fac_phi (p, n)
  int nl[2] = {0, 0};
  bool fac_list_ok = false;
  std::string fac_reason;
  double *d_Uf = nullptr, *d_Vf = nullptr, *d_UfT = nullptr, *d_VfT = nullptr;
  // Schur route of the factored Pfaffian (set when every species block of
  // every state has cond(Z_s) <= PFS_COND_MAX): (-Z_s)^-1 [nstates][2][h*h]
  // and Pf(-cscale Z_s) [nstates][2]
  double *d_binv = nullptr, *d_pfb = nullptr;
  int *d_pfs_ok = nullptr;             // [nstates]: 1 if both species blocks of the state pass
  bool pfs_any = false, pfs_all = false;  // some / every state passes
  bool schur_allowed = false;          // the caller asked for it (pnamp_set_schur_pfaffian)
  double *d_mtab = nullptr;         // [3][h]: m, <m+1|J_+|m>, j(j+1) per species-local state
  double2 *d_fout = nullptr, *d_frho = nullptr, *d_sbuf = nullptr;
  int *d_fstatus = nullptr;
  int2 *d_pn_rows = nullptr;
  int *d_pn_blk = nullptr, *d_pn_ycol = nullptr;
  double *d_pn_w = nullptr;
  int pn_ncol = 0;
  size_t fac_cap = 0;  // pairs the factored scratch holds
  // captured launch sequence of the factored form, valid for (npairs, zval,
  // nval) and the current buffers; dropped whenever a buffer is reallocated
  hipGraphExec_t fac_graph = nullptr;
  int state_cap = 0, sep_cap = 0;   // states the state buffers (and the species blocks) hold
  int graph_captures = 0, graph_replays = 0, graph_nodes = 0;
  FacTune tune;                     // occupancy variants and chunking of the FP64 factored form
  int g_np = -1, g_z = 0, g_n = 0, g_form = -1, g_zn = 0;
  // mixed forms: the dense V-contraction matrices (W_pp, W_nn, Q_pp, Q_nn,
  // W_pn; kp x kp row-major FP64 on the host), their fragment-ordered device
  // copies in the precision of va_mix, and the per-chunk GEMM operands
  int kp = 0;
  std::vector<double> vmat[NVPROB];
  void *d_va[NVPROB] = {}, *d_va_lo[NVPROB] = {};
  int va_mix = -1;
  void *d_brho[2] = {}, *d_brho_lo[2] = {}, *d_ba10[2] = {}, *d_ba10_lo[2] = {};
  double2 *d_a01 = nullptr;
  float *d_gpn = nullptr;
  size_t mix_cap = 0;  // pairs the mixed-form operands hold, in the precision mix_cap_mix
  int mix_cap_mix = -1;
};

#define HIPCHK(ctx, x)                                                       \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      (ctx)->err = std::string(#x) + ": " + hipGetErrorString(e_);           \
      return PNAMP_E_DEVICE;                                                 \
    }                                                                        \
  } while (0)

static int fail(pnamp_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

static void drop_graph(pnamp_ctx *c) {
  if (c && c->fac_graph) {
    (void)hipGraphExecDestroy(c->fac_graph);
    c->fac_graph = nullptr;
    c->g_np = -1;
    c->g_form = -1;
  }
}

static void free_void(pnamp_ctx *c, void **p) {
  drop_graph(c);
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

// Fragment-ordered copy of a kp x kp row-major matrix for k_vgemm (A operand).
static int upload_vmat(pnamp_ctx *c, int idx, int mix) {
  const int kp = c->kp;
  const std::vector<double> &A = c->vmat[idx];
  free_void(c, &c->d_va[idx]);
  free_void(c, &c->d_va_lo[idx]);
  if (mix == MIX_FP32) {
    const int kq_n = kp / 8;
    std::vector<float> sw((size_t)kp * kp);
    for (int rb = 0; rb < kp / 32; ++rb)
      for (int kq = 0; kq < kq_n; ++kq)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 4; ++e)
            sw[(((size_t)kq * (kp / 32) + rb) * 64 + l) * 4 + e] =
                (float)A[(size_t)(32 * rb + (l & 31)) * kp + (8 * kq + 2 * e + (l >> 5))];
    HIPCHK(c, hipMalloc(&c->d_va[idx], sw.size() * sizeof(float)));
    HIPCHK(c, hipMemcpy(c->d_va[idx], sw.data(), sw.size() * sizeof(float), hipMemcpyHostToDevice));
  } else {
    const int kb_n = kp / 16;
    std::vector<unsigned short> hi((size_t)kp * kp), lo((size_t)kp * kp);
    for (int rb = 0; rb < kp / 32; ++rb)
      for (int kb = 0; kb < kb_n; ++kb)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e) {
            const size_t i = (((size_t)kb * (kp / 32) + rb) * 64 + l) * 8 + e;
            const double v = A[(size_t)(32 * rb + (l & 31)) * kp + (16 * kb + 8 * (l >> 5) + e)];
            hi[i] = bf16_rne((float)v);
            lo[i] = bf16_rne((float)(v - (double)bf16_float(hi[i])));
          }
    HIPCHK(c, hipMalloc(&c->d_va[idx], hi.size() * 2));
    HIPCHK(c, hipMalloc(&c->d_va_lo[idx], lo.size() * 2));
    HIPCHK(c, hipMemcpy(c->d_va[idx], hi.data(), hi.size() * 2, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_va_lo[idx], lo.data(), lo.size() * 2, hipMemcpyHostToDevice));
  }
  return 0;
}

template <class T>
static int dalloc(pnamp_ctx *c, T **p, size_t count) {
  drop_graph(c);
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  HIPCHK(c, hipMalloc((void **)p, count * sizeof(T)));
  return 0;
}

template <class T>
static int dupload(pnamp_ctx *c, T **p, const T *src, size_t count) {
  int r = dalloc(c, p, count);
  if (r) return r;
  if (count) HIPCHK(c, hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int pnamp_create(const pnamp_basis *basis, const pnamp_interaction *inter,
                            const pnamp_mesh *mesh, int device, pnamp_ctx **out) {
  if (!out) return PNAMP_E_ARG;
  *out = nullptr;
  if (!basis || !inter || !mesh) return PNAMP_E_ARG;
  pnamp_ctx *c = new pnamp_ctx();
  auto bail = [&](int code, const std::string &m) {
    c->err = m;
    *out = c;  // keep the context so the caller can read the message
    return code;
  };
  const int n = basis->n;
  if (n <= 0 || (n & 1)) return bail(PNAMP_E_ARG, "n must be positive and even");
  if (n > NMAX || !plain_supported(n)) return bail(PNAMP_E_SIZE, "n exceeds PNAMP_NMAX");
  const int h = n / 2;
  for (int i = 0; i < n; ++i)
    if (basis->t[i] != (i < h ? -1 : 1)) return bail(PNAMP_E_ARG, "basis must list protons (t=-1) then neutrons (t=+1)");
  for (int i = 0; i < h; ++i)
    if (basis->nq[i] != basis->nq[i + h] || basis->l[i] != basis->l[i + h] || basis->twoj[i] != basis->twoj[i + h] ||
        basis->twom[i] != basis->twom[i + h])
      return bail(PNAMP_E_ARG, "proton and neutron halves must have the same orbits");
  // rotation blocks: contiguous runs of equal (n, l, j), m = j..-j
  std::vector<int> off(h), sz(h), base(h), pa, pb;
  for (int i = 0; i < h;) {
    const int tj = basis->twoj[i];
    const int s = tj + 1;
    if (i + s > h) return bail(PNAMP_E_ARG, "incomplete orbit block");
    for (int k = 0; k < s; ++k) {
      if (basis->twoj[i + k] != tj || basis->nq[i + k] != basis->nq[i] || basis->l[i + k] != basis->l[i] ||
          basis->twom[i + k] != tj - 2 * k)
        return bail(PNAMP_E_ARG, "orbit block must run m = j .. -j");
      off[i + k] = i;
      sz[i + k] = s;
      base[i + k] = (int)pa.size();
    }
    for (int bcol = 0; bcol < s; ++bcol)
      for (int arow = 0; arow < s; ++arow) {
        pa.push_back(i + arow);
        pb.push_back(i + bcol);
      }
    i += s;
  }
  if ((int)pa.size() > RPACK_MAX) return bail(PNAMP_E_SIZE, "rotation blocks exceed the packed LDS table");
  if (mesh->nalp < 1 || mesh->nbet < 1 || mesh->ngam < 1 || mesh->nphi < 1 || mesh->jmax < 0)
    return bail(PNAMP_E_ARG, "bad mesh sizes");
  if (mesh->jmax > JMAX_MAX) return bail(PNAMP_E_SIZE, "jmax too large");
  if ((2 * (size_t)mesh->jmax + 1) * (mesh->nalp + mesh->ngam) * sizeof(double2) > 150 * 1024)
    return bail(PNAMP_E_SIZE, "alpha and gamma meshes too large for the projection phase tables");
  if (mesh->fold != PNAMP_FOLD_PLAIN && mesh->fold != PNAMP_FOLD_LITERAL) return bail(PNAMP_E_ARG, "bad fold");
  if (inter->nme < 0 || inter->nme > (1 << 30)) return bail(PNAMP_E_ARG, "bad interaction size");
  c->device = device;
  HIPCHK(c, hipSetDevice(device));
  c->scratch = noalloc_scratch_ptr(device, c->err);
  if (!c->scratch) return bail(PNAMP_E_DEVICE, c->err);
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto &e : c->ev) HIPCHK(c, hipEventCreate(&e));
  HIPCHK(c, plain_prepare(n, false));
  HIPCHK(c, plain_prepare(n, true));
  HIPCHK(c, vgemm_attr());
  HIPCHK(c, hipFuncSetAttribute((const void *)k_state_prep, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * (2 * (size_t)n * n + n))));
  double fact[40];
  fact[0] = 1.0;
  for (int i = 1; i < 40; ++i) fact[i] = fact[i - 1] * i;
  HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_fact), fact, sizeof(fact)));
  c->n = n;
  c->h = h;
  c->npack = (int)pa.size();
  c->twom_h.assign(basis->twom, basis->twom + h);
  c->twoj_h.assign(basis->twoj, basis->twoj + h);
  c->l_h.assign(basis->l, basis->l + h);
  c->nq_h.assign(basis->nq, basis->nq + h);
  c->nalp = mesh->nalp;
  c->nbet = mesh->nbet;
  c->ngam = mesh->ngam;
  c->nphi = mesh->nphi;
  c->jmax = mesh->jmax;
  c->fold = mesh->fold;
  c->NE = mesh->nalp * mesh->nbet * mesh->ngam;
  c->npj = 0;
  for (int J = 0; J <= c->jmax; ++J) c->npj += (2 * J + 1) * (2 * J + 1);
  c->cscale = mesh->cscale;
  c->vac_scale = std::pow(10.0, (double)(-(n * mesh->iscale) / 2));
  const long aaa = (long)n * (n + 1) / 2;
  c->snpfaf = (aaa % 2) ? -1.0 : 1.0;
  c->e0 = inter->e0;
  c->nme = (int)inter->nme;
  int r = 0;
  std::vector<double> bet(mesh->bet, mesh->bet + mesh->nbet), wbet(mesh->wbet, mesh->wbet + mesh->nbet);
  if ((r = dupload(c, &c->d_alp, mesh->alp, c->nalp)) || (r = dupload(c, &c->d_walp, mesh->walp, c->nalp)) ||
      (r = dupload(c, &c->d_bet, bet.data(), c->nbet)) || (r = dupload(c, &c->d_wbet, wbet.data(), c->nbet)) ||
      (r = dupload(c, &c->d_gam, mesh->gam, c->ngam)) || (r = dupload(c, &c->d_wgam, mesh->wgam, c->ngam)))
    return bail(r, c->err);
  {
    // ladder-operator table of the factored form, per species-local state a
    // (m = j .. -j inside an orbit, so a - 1 is the m + 1 partner of a)
    std::vector<double> mt(3 * (size_t)h, 0.0);
    for (int a = 0; a < h; ++a) {
      const double j = basis->twoj[a] / 2.0, m = basis->twom[a] / 2.0;
      mt[a] = m;
      mt[h + a] = (a > off[a]) ? std::sqrt(j * (j + 1) - m * (m + 1)) : 0.0;
      mt[2 * h + a] = j * (j + 1);
    }
    if ((r = dupload(c, &c->d_mtab, mt.data(), mt.size()))) return bail(r, c->err);
    // per species and element (a, b) of rho, e = a + b h: t1(b, a), m_a m_b,
    // c_a c_b (0 unless a, b > 0), d_ab, m_a d_ab, j_a(j_a+1) d_ab,
    // c_a d_{a,b+1}, c_b d_{b,a+1} (the kinetic energy and <J^2> loop of k_fac_phi)
    std::vector<double> jt(2 * (size_t)h * h * 8, 0.0);
    for (int sp = 0; sp < 2; ++sp)
      for (int b = 0; b < h; ++b)
        for (int a = 0; a < h; ++a) {
          double *q = jt.data() + ((size_t)sp * h * h + a + (size_t)b * h) * 8;
          const double ma = mt[a], mb = mt[b], ca = mt[h + a], cb = mt[h + b];
          q[0] = inter->t1b[(size_t)(sp * h + b) + (size_t)(sp * h + a) * n];
          q[1] = ma * mb;
          q[2] = (a > 0 && b > 0) ? ca * cb : 0.0;
          q[3] = (a == b) ? 1.0 : 0.0;
          q[4] = (a == b) ? ma : 0.0;
          q[5] = (a == b) ? mt[2 * h + a] : 0.0;
          q[6] = (a == b + 1) ? ca : 0.0;
          q[7] = (b == a + 1) ? cb : 0.0;
        }
    if ((r = dupload(c, &c->d_jtab, jt.data(), jt.size()))) return bail(r, c->err);
  }
  if ((r = dupload(c, &c->d_pk_a, pa.data(), pa.size())) || (r = dupload(c, &c->d_pk_b, pb.data(), pb.size())) ||
      (r = dupload(c, &c->d_blk_off, off.data(), h)) || (r = dupload(c, &c->d_blk_sz, sz.data(), h)) ||
      (r = dupload(c, &c->d_blk_base, base.data(), h)) ||
      (r = dupload(c, &c->d_twoj, (const int *)basis->twoj, h)) ||
      (r = dupload(c, &c->d_twom, (const int *)basis->twom, h)))
    return bail(r, c->err);
  // plain-form energy list: each entry (a b | V | c d) with its pairs in
  // increasing order (<ba|V|cd> = -<ab|V|cd>), as the term
  //   V [rho(c,a) rho(d,b) - rho(d,a) rho(c,b) + A01(a,b) A10(c,d) / 4]
  // and, when a + b != c + d, its (cd)(ab) partner; an entry with a = b or
  // c = d adds exactly zero to the eight-term sum of HFB_Energy_COMPLEX and is
  // dropped.  Operands by LDS offset (rho at leading dimension n + 1, then
  // the A-block: A01(a,b) at a + b n, A10(c,d) / 4 at d + c n): bytes when
  // they fit 16 bits, else elements.
  {
    const unsigned ldr = n + 1, abase = n * ldr;
    c->pl_bytes = (16u * (abase + (unsigned)n * n) <= 65535u) ? 1 : 0;
    const unsigned sc = c->pl_bytes ? 16u : 1u;
    auto rb = [&](int row, int col) { return sc * ((unsigned)row + (unsigned)col * ldr); };
    auto ab = [&](int row, int col) { return sc * (abase + (unsigned)row + (unsigned)col * n); };
    std::vector<uint4> lo;
    std::vector<double> lm;
    auto term = [&](int i, int j, int k, int l, double me) {
      lo.push_back(make_uint4(rb(k, i) | (rb(l, j) << 16), rb(l, i) | (rb(k, j) << 16), ab(i, j) | (ab(l, k) << 16), 0u));
      lm.push_back(me);
    };
    for (int64_t q = 0; q < inter->nme; ++q) {
      int a = inter->ka[q] - 1, b = inter->kb[q] - 1, cc = inter->kc[q] - 1, d = inter->kd[q] - 1;
      if (a < 0 || b < 0 || cc < 0 || d < 0 || a >= n || b >= n || cc >= n || d >= n)
        return bail(PNAMP_E_ARG, "interaction index out of range");
      double me = inter->me2b[q];
      if (a == b || cc == d) continue;
      if (a > b) { std::swap(a, b); me = -me; }
      if (cc > d) { std::swap(cc, d); me = -me; }
      term(a, b, cc, d, me);
      if (cc + d != a + b) term(cc, d, a, b, me);
    }
    while (lo.size() % 256 || lo.empty()) {  // zero terms read rho(0,0) and add nothing
      lo.push_back(make_uint4(0u, 0u, 0u, 0u));
      lm.push_back(0.0);
    }
    c->pl_n = (int)lo.size();
    if ((r = dupload(c, &c->d_plofs, lo.data(), lo.size())) || (r = dupload(c, &c->d_plme, lm.data(), lm.size())) ||
        (r = dupload(c, &c->d_t1b, inter->t1b, (size_t)n * n)))
      return bail(r, c->err);
  }
  // species lists of the factored form: pp and nn with species-local indices,
  // pn (a, c proton; b, d neutron) with the (ab) <-> (cd) partner flag in pad
  {
    std::vector<VEntry> ls[3];
    c->fac_list_ok = true;
    for (int64_t q = 0; q < inter->nme; ++q) {
      int a = inter->ka[q] - 1, b = inter->kb[q] - 1, cc = inter->kc[q] - 1, d = inter->kd[q] - 1;
      int ta = a >= h, tb = b >= h, tc = cc >= h, td = d >= h;
      VEntry e{inter->me2b[q], 0u, 0u};
      if (ta + tb != tc + td) {  // changes the charge: the plain form takes the list as it is
        c->fac_list_ok = false;
        c->fac_reason = "factored and mixed forms need list entries that conserve the charge";
        continue;
      }
      // the same matrix element with each pair in canonical order (proton
      // first, else the smaller index first): <ba|V|cd> = -<ab|V|cd>; the
      // index sums a + b, c + d (the partner rule) are unchanged
      if (ta > tb || (ta == tb && a > b)) {
        std::swap(a, b);
        std::swap(ta, tb);
        e.me = -e.me;
      }
      if (tc > td || (tc == td && cc > d)) {
        std::swap(cc, d);
        std::swap(tc, td);
        e.me = -e.me;
      }
      if (!ta && !tb && !tc && !td) {
        e.idx = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)cc << 16) | ((uint32_t)d << 24);
        ls[0].push_back(e);
      } else if (ta && tb && tc && td) {
        e.idx = (uint32_t)(a - h) | ((uint32_t)(b - h) << 8) | ((uint32_t)(cc - h) << 16) | ((uint32_t)(d - h) << 24);
        ls[1].push_back(e);
      } else if (!ta && tb && !tc && td) {
        e.idx = (uint32_t)a | ((uint32_t)(b - h) << 8) | ((uint32_t)cc << 16) | ((uint32_t)(d - h) << 24);
        e.pad = (a + b == cc + d) ? 0u : 1u;
        ls[2].push_back(e);
      }
    }
    for (int k = 0; k < 3; ++k) c->nme_s[k] = (int)ls[k].size();
    // pp and nn lists for k_fac_phi: one term per entry, the (kl)(ij) partner
    // of an entry with a + b != c + d as an entry of its own, LDS byte
    // offsets precomputed (rho at ld h + 1, then the A-block at ld h)
    for (int k = 0; k < 2; ++k) {
      std::vector<uint4> lo, lw;
      const unsigned hp = h + 1, msz = h * (h + 1);
      auto rb = [&](int row, int col) { return (unsigned)(row + col * hp) * 16u; };
      auto ab = [&](int row, int col) { return (msz + (unsigned)(row + col * h)) * 16u; };
      auto term = [&](int i, int j, int kk, int l, double me) {
        uint64_t mb;
        std::memcpy(&mb, &me, 8);
        lo.push_back(make_uint4(rb(kk, i), rb(l, j), rb(l, i), rb(kk, j)));
        lw.push_back(make_uint4(ab(i, j), ab(l, kk), (unsigned)mb, (unsigned)(mb >> 32)));
      };
      for (const VEntry &e : ls[k]) {
        const int i = e.idx & 255, j = (e.idx >> 8) & 255, kk = (e.idx >> 16) & 255, l = e.idx >> 24;
        if (!(i < j && kk < l)) {  // a = b or c = d: the A-block holds index pairs in increasing order only
          c->fac_list_ok = false;
          c->fac_reason = "factored and mixed forms need list entries with a != b and c != d";
          break;
        }
        term(i, j, kk, l, e.me);
        if (kk + l != i + j) term(kk, l, i, j, e.me);
      }
      while (lo.size() % 256) {  // zero terms (offset 0, matrix element 0)
        lo.push_back(make_uint4(0u, 0u, 0u, 0u));
        lw.push_back(make_uint4(0u, 0u, 0u, 0u));
      }
      c->nl[k] = (int)lo.size();
      lo.insert(lo.end(), lw.begin(), lw.end());
      if ((r = dupload(c, &c->d_lofs[k], lo.data(), lo.size()))) return bail(r, c->err);
    }
    // W_pn by rows x = c + a h, columns y = d + b h (species-local indices)
    const int hh = h * h;
    std::vector<std::vector<std::pair<int, double>>> rows(hh);
    auto add = [&](int x, int y, double v) {
      for (auto &e : rows[x])
        if (e.first == y) { e.second += v; return; }
      rows[x].push_back({y, v});
    };
    for (const VEntry &e : ls[2]) {
      const int a = e.idx & 255, b = (e.idx >> 8) & 255, cc = (e.idx >> 16) & 255, d = e.idx >> 24;
      add(cc + a * h, d + b * h, e.me);
      if (e.pad) add(a + cc * h, b + d * h, e.me);
    }
    std::vector<int> wptr(hh + 1, 0), wcol;
    std::vector<double> wval;
    for (int x = 0; x < hh; ++x) {
      std::sort(rows[x].begin(), rows[x].end());
      for (auto &e : rows[x]) { wcol.push_back(e.first); wval.push_back(e.second); }
      wptr[x + 1] = (int)wcol.size();
    }
    // dense Delta-m blocks of W_pn for k_fac_combine: row x = c + a h has
    // Delta m = m_c - m_a, column y = d + b h has m_d - m_b, and a non-zero
    // needs Delta m(x) = -Delta m(y)
    {
      auto dm = [&](int idx) { return basis->twom[idx % h] - basis->twom[idx / h]; };
      std::vector<int> dvals;
      for (int x = 0; x < hh; ++x) dvals.push_back(dm(x));
      std::sort(dvals.begin(), dvals.end());
      dvals.erase(std::unique(dvals.begin(), dvals.end()), dvals.end());
      std::vector<int2> prow;
      std::vector<int> blk, ycol;
      std::vector<double> wblk;
      for (size_t bi = 0; bi < dvals.size(); ++bi) {
        const int D = dvals[bi];
        std::vector<int> xs, ys;
        for (int x = 0; x < hh; ++x) if (dm(x) == D) xs.push_back(x);
        for (int y = 0; y < hh; ++y) if (dm(y) == -D) ys.push_back(y);
        blk.insert(blk.end(), {(int)prow.size(), (int)xs.size(), (int)ycol.size(), (int)ys.size(), (int)wblk.size()});
        const size_t w0 = wblk.size();
        wblk.resize(w0 + xs.size() * ys.size(), 0.0);
        for (size_t i = 0; i < xs.size(); ++i) {
          for (const auto &e : rows[xs[i]]) {
            const auto it = std::lower_bound(ys.begin(), ys.end(), e.first);
            if (it == ys.end() || *it != e.first) {  // the plain form takes the list as it is
              c->fac_list_ok = false;
              c->fac_reason = "factored and mixed forms need proton-neutron list entries that conserve M";
              continue;
            }
            wblk[w0 + (size_t)(it - ys.begin()) * xs.size() + i] = e.second;
          }
          prow.push_back(make_int2(xs[i], (int)bi));
        }
        ycol.insert(ycol.end(), ys.begin(), ys.end());
      }
      c->pn_ncol = (int)ycol.size();
      if ((r = dupload(c, &c->d_pn_rows, prow.data(), prow.size())) || (r = dupload(c, &c->d_pn_blk, blk.data(), blk.size())) ||
          (r = dupload(c, &c->d_pn_ycol, ycol.data(), ycol.size())) || (r = dupload(c, &c->d_pn_w, wblk.data(), wblk.size())))
        return bail(r, c->err);
    }
    // dense forms of the mixed-precision contraction (rows x, columns y, x and
    // y = species-local pair index c + a h of rho(c, a)): for each species entry
    // (i j | V | k l) with d = [k + l != i + j]
    //   W[(k,i),(l,j)] += V, W[(l,i),(k,j)] -= V, and with d W[(i,k),(j,l)] += V, W[(j,k),(i,l)] -= V,
    //   Q[(i,j),(k,l)] += V/4, and with d Q[(k,l),(i,j)] += V/4,
    // so that rho^T W rho + A01^T Q A10 is the regrouped eight-term sum; W_pn from the rows above
    const int kp = (hh + 31) / 32 * 32;
    c->kp = kp;
    for (auto &m : c->vmat) m.assign((size_t)kp * kp, 0.0);
    auto at = [&](int x, int y) { return (size_t)x * kp + y; };
    for (int sp = 0; sp < 2; ++sp) {
      std::vector<double> &W = c->vmat[sp], &Q = c->vmat[2 + sp];
      for (const VEntry &e : ls[sp]) {
        const int i = e.idx & 255, j = (e.idx >> 8) & 255, k = (e.idx >> 16) & 255, l = e.idx >> 24;
        const double v = e.me;
        W[at(k + i * h, l + j * h)] += v;
        W[at(l + i * h, k + j * h)] -= v;
        Q[at(i + j * h, k + l * h)] += 0.25 * v;
        if (k + l != i + j) {
          W[at(i + k * h, j + l * h)] += v;
          W[at(j + k * h, i + l * h)] -= v;
          Q[at(k + l * h, i + j * h)] += 0.25 * v;
        }
      }
    }
    for (int x = 0; x < hh; ++x)
      for (int q = wptr[x]; q < wptr[x + 1]; ++q) c->vmat[4][at(x, wcol[q])] = wval[q];
  }
  if ((r = dalloc(c, &c->d_rtab, (size_t)c->nbet * c->npack)) || (r = dalloc(c, &c->d_ptab, (size_t)c->nbet * 2 * c->npj)))
    return bail(r, c->err);
  TableParams tp{c->nbet, h, c->npack, c->jmax, c->npj, c->d_bet, c->d_pk_a, c->d_pk_b, c->d_twoj, c->d_twom,
                 c->d_rtab, c->d_ptab};
  hipLaunchKernelGGL(k_dtables, dim3(c->nbet), dim3(256), 0, c->stream, tp);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = c;
  return PNAMP_OK;
}

// Cyclic Jacobi for a small real symmetric matrix (host): A = Q diag(w) Q^T,
// A and Q column-major n x n.
static void host_jacobi(int n, std::vector<double> A, std::vector<double> &Q, std::vector<double> &w) {
  Q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Q[i + (size_t)i * n] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double x = A[i + (size_t)j * n] * A[i + (size_t)j * n];
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int r = p + 1; r < n; ++r) {
        const double apr = A[p + (size_t)r * n];
        if (std::fabs(apr) < 1e-300) continue;
        const double th = (A[r + (size_t)r * n] - A[p + (size_t)p * n]) / (2.0 * apr);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {  // columns p, r
          const double x = A[k + (size_t)p * n], y = A[k + (size_t)r * n];
          A[k + (size_t)p * n] = cs * x - sn * y;
          A[k + (size_t)r * n] = sn * x + cs * y;
        }
        for (int k = 0; k < n; ++k) {  // rows p, r
          const double x = A[p + (size_t)k * n], y = A[r + (size_t)k * n];
          A[p + (size_t)k * n] = cs * x - sn * y;
          A[r + (size_t)k * n] = sn * x + cs * y;
        }
        for (int k = 0; k < n; ++k) {
          const double x = Q[k + (size_t)p * n], y = Q[k + (size_t)r * n];
          Q[k + (size_t)p * n] = cs * x - sn * y;
          Q[k + (size_t)r * n] = sn * x + cs * y;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i + (size_t)i * n];
}

// Largest cond_1(Z_s) = |Z_s|_1 |Z_s^-1|_1 of the Schur route of the factored
// Pfaffian: its Pfaffians carry up to about cond(Z_s) eps of relative rounding
// (at most 6.1e-11 over random rotations of states with cond_1 up to 5.6e5).
constexpr double PFS_COND_MAX = 1.0e6;

// Inverse of a real m x m matrix (column-major) by Gauss-Jordan with partial
// pivoting; false if a pivot is exactly zero.
static bool host_inverse(int m, std::vector<double> A, std::vector<double> &X) {
  X.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) X[i + (size_t)i * m] = 1.0;
  for (int k = 0; k < m; ++k) {
    int p = k;
    for (int i = k + 1; i < m; ++i)
      if (std::fabs(A[i + (size_t)k * m]) > std::fabs(A[p + (size_t)k * m])) p = i;
    if (A[p + (size_t)k * m] == 0.0) return false;
    for (int j = 0; j < m; ++j) {
      std::swap(A[k + (size_t)j * m], A[p + (size_t)j * m]);
      std::swap(X[k + (size_t)j * m], X[p + (size_t)j * m]);
    }
    const double inv = 1.0 / A[k + (size_t)k * m];
    for (int j = 0; j < m; ++j) {
      A[k + (size_t)j * m] *= inv;
      X[k + (size_t)j * m] *= inv;
    }
    for (int i = 0; i < m; ++i) {
      if (i == k) continue;
      const double f = A[i + (size_t)k * m];
      if (f == 0.0) continue;
      for (int j = 0; j < m; ++j) {
        A[i + (size_t)j * m] -= f * A[k + (size_t)j * m];
        X[i + (size_t)j * m] -= f * X[k + (size_t)j * m];
      }
    }
  }
  return true;
}

// Pfaffian of a real skew m x m matrix (column-major, m even) by Parlett-Reid
// with the pivot of largest modulus in row k.
static double host_pfaffian(int m, std::vector<double> A) {
  double pf = 1.0;
  for (int k = 0; k + 1 < m; k += 2) {
    int c = k + 1;
    for (int j = k + 2; j < m; ++j)
      if (std::fabs(A[k + (size_t)j * m]) > std::fabs(A[k + (size_t)c * m])) c = j;
    if (c != k + 1) {
      for (int i = 0; i < m; ++i) std::swap(A[i + (size_t)(k + 1) * m], A[i + (size_t)c * m]);
      for (int j = 0; j < m; ++j) std::swap(A[(k + 1) + (size_t)j * m], A[c + (size_t)j * m]);
      pf = -pf;
    }
    const double s = A[k + (size_t)(k + 1) * m];
    if (s == 0.0) return 0.0;
    pf *= s;
    // S(i, j) += (S(i, k+1) S(j, k) - S(i, k) S(j, k+1)) / s for i, j > k + 1
    for (int j = k + 2; j < m; ++j)
      for (int i = k + 2; i < m; ++i)
        A[i + (size_t)j * m] += (A[i + (size_t)(k + 1) * m] * A[j + (size_t)k * m] -
                                 A[i + (size_t)k * m] * A[j + (size_t)(k + 1) * m]) / s;
  }
  return pf;
}

static double host_norm1(int m, const std::vector<double> &A) {
  double r = 0.0;
  for (int j = 0; j < m; ++j) {
    double col = 0.0;
    for (int i = 0; i < m; ++i) col += std::fabs(A[i + (size_t)j * m]);
    r = std::max(r, col);
  }
  return r;
}

extern "C" int pnamp_set_states(pnamp_ctx *c, int nstates, const double *U, const double *V) {
  if (!c || !U || !V || nstates <= 0) return PNAMP_E_ARG;
  const size_t nn = (size_t)c->n * c->n;
  int r = 0;
  c->separable = false;
  if (nstates <= c->state_cap) {
    // same buffers: a captured launch graph (which holds their addresses) stays valid
    HIPCHK(c, hipMemcpy(c->d_U, U, sizeof(double) * nn * nstates, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_V, V, sizeof(double) * nn * nstates, hipMemcpyHostToDevice));
  } else {
    c->state_cap = 0;
    if ((r = dupload(c, &c->d_U, U, nn * nstates)) || (r = dupload(c, &c->d_V, V, nn * nstates)) ||
        (r = dalloc(c, &c->d_Z, nn * nstates)) || (r = dalloc(c, &c->d_det, (size_t)nstates)) ||
        (r = dalloc(c, &c->d_state_status, (size_t)nstates)))
      return r;
    c->state_cap = nstates;
  }
  PrepParams pp{c->n, c->d_U, c->d_V, c->d_Z, c->d_det, c->d_state_status};
  hipLaunchKernelGGL(k_state_prep, dim3(nstates), dim3(256), sizeof(double) * (2 * nn + c->n), c->stream, pp);
  HIPCHK(c, hipGetLastError());
  std::vector<int> st(nstates);
  HIPCHK(c, hipMemcpyAsync(st.data(), c->d_state_status, sizeof(int) * nstates, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nstates = nstates;
  for (int q = 0; q < nstates; ++q)
    if (st[q]) {
      c->nstates = 0;
      return fail(c, st[q], "U of state " + std::to_string(q) + " is singular (zbareover ZGETRF INFO)");
    }
  // Species separability, decided on Z = V U^-1 (the state itself, whatever
  // the quasiparticle basis): no proton-neutron block.  Then each species is
  // the Bogoliubov block U_s = (1 + Z_s^T Z_s)^(-1/2), V_s = Z_s U_s, which
  // has V_s U_s^-1 = Z_s and U_s^T U_s + V_s^T V_s = 1; the kernels depend
  // only on Z_s (transition densities and overlaps are invariant under
  // U -> U C, V -> V C), so the factored form may use these blocks.
  const int n = c->n, h = c->h;
  std::vector<double> Zh(nn * nstates);
  HIPCHK(c, hipMemcpy(Zh.data(), c->d_Z, sizeof(double) * Zh.size(), hipMemcpyDeviceToHost));
  std::vector<double> Uf((size_t)nstates * 2 * h * h), Vf((size_t)nstates * 2 * h * h);
  std::vector<double> Binv((size_t)nstates * 2 * h * h), pfb((size_t)nstates * 2);
  std::vector<int> pfs_ok(nstates, 1);
  bool sep = true;
  for (int q = 0; q < nstates && sep; ++q) {
    const double *Zq = Zh.data() + (size_t)q * nn;
    double scale = 0.0, off = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double v = std::fabs(Zq[i + (size_t)j * n]);
        scale = std::max(scale, v);
        if ((i < h) != (j < h)) off = std::max(off, v);
      }
    if (off > 1e-11 * std::max(scale, 1.0)) {  // rounding of U^-1 leaves ~1e-14 in a separable state
      sep = false;
      break;
    }
    for (int spc = 0; spc < 2; ++spc) {
      // M = Z_s^T Z_s (symmetric, non-negative): M = Q diag(w) Q^T by Jacobi,
      // U_s = Q diag((1 + w)^(-1/2)) Q^T
      std::vector<double> M((size_t)h * h, 0.0), Q, w;
      for (int j = 0; j < h; ++j)
        for (int i = 0; i < h; ++i) {
          double acc = 0.0;
          for (int k = 0; k < h; ++k)
            acc += Zq[(spc * h + k) + (size_t)(spc * h + i) * n] * Zq[(spc * h + k) + (size_t)(spc * h + j) * n];
          M[i + (size_t)j * h] = acc;
        }
      host_jacobi(h, M, Q, w);
      double *Us = Uf.data() + ((size_t)q * 2 + spc) * h * h, *Vs = Vf.data() + ((size_t)q * 2 + spc) * h * h;
      for (int j = 0; j < h; ++j)
        for (int i = 0; i < h; ++i) {
          double acc = 0.0;
          for (int k = 0; k < h; ++k) acc += Q[i + (size_t)k * h] * Q[j + (size_t)k * h] / std::sqrt(1.0 + w[k]);
          Us[i + (size_t)j * h] = acc;
        }
      for (int j = 0; j < h; ++j)
        for (int i = 0; i < h; ++i) {
          double acc = 0.0;
          for (int k = 0; k < h; ++k) acc += Zq[(spc * h + i) + (size_t)(spc * h + k) * n] * Us[k + (size_t)j * h];
          Vs[i + (size_t)j * h] = acc;
        }
      // Schur route: B = -Z_s from its strict upper triangle (the entries the
      // Pfaffian kernel reads), B^-1 and Pf(cscale B)
      if (pfs_ok[q]) {
        std::vector<double> B((size_t)h * h, 0.0), X;
        for (int j = 0; j < h; ++j)
          for (int i = 0; i < j; ++i) {
            const double z = Zq[(spc * h + i) + (size_t)(spc * h + j) * n];
            B[i + (size_t)j * h] = -z;
            B[j + (size_t)i * h] = z;
          }
        if (!host_inverse(h, B, X) || !(host_norm1(h, B) * host_norm1(h, X) <= PFS_COND_MAX)) {
          pfs_ok[q] = 0;
        } else {
          std::copy(X.begin(), X.end(), Binv.begin() + ((size_t)q * 2 + spc) * h * h);
          for (double &x : B) x *= c->cscale;
          pfb[(size_t)q * 2 + spc] = host_pfaffian(h, B);
        }
      }
    }
  }
  c->separable = sep;
  bool any = false, all = sep;
  for (int q = 0; q < nstates; ++q) {
    any = any || (sep && pfs_ok[q]);
    all = all && pfs_ok[q];
  }
  if (any != c->pfs_any || all != c->pfs_all) drop_graph(c);  // the captured launches name the Pfaffian kernels
  c->pfs_any = any;
  c->pfs_all = all;
  if (sep) {
    int r2 = 0;
    // row-major copies (U_s(j, k) at j h + k) for the wave-uniform reads of k_fac_phi
    std::vector<double> UfT(Uf.size()), VfT(Vf.size());
    for (size_t b = 0; b < (size_t)nstates * 2; ++b)
      for (int j = 0; j < h; ++j)
        for (int k = 0; k < h; ++k) {
          UfT[b * h * h + (size_t)j * h + k] = Uf[b * h * h + j + (size_t)k * h];
          VfT[b * h * h + (size_t)j * h + k] = Vf[b * h * h + j + (size_t)k * h];
        }
    if (nstates <= c->sep_cap) {
      const size_t b = sizeof(double) * Uf.size();
      HIPCHK(c, hipMemcpy(c->d_Uf, Uf.data(), b, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->d_Vf, Vf.data(), b, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->d_UfT, UfT.data(), b, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->d_VfT, VfT.data(), b, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->d_binv, Binv.data(), b, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->d_pfb, pfb.data(), sizeof(double) * pfb.size(), hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->d_pfs_ok, pfs_ok.data(), sizeof(int) * pfs_ok.size(), hipMemcpyHostToDevice));
    } else {
      c->sep_cap = 0;
      if ((r2 = dupload(c, &c->d_Uf, Uf.data(), Uf.size())) || (r2 = dupload(c, &c->d_Vf, Vf.data(), Vf.size())) ||
          (r2 = dupload(c, &c->d_UfT, UfT.data(), UfT.size())) || (r2 = dupload(c, &c->d_VfT, VfT.data(), VfT.size())) ||
          (r2 = dupload(c, &c->d_binv, Binv.data(), Binv.size())) || (r2 = dupload(c, &c->d_pfb, pfb.data(), pfb.size())) ||
          (r2 = dupload(c, &c->d_pfs_ok, pfs_ok.data(), pfs_ok.size())))
        return r2;
      c->sep_cap = nstates;
    }
  }
  return PNAMP_OK;
}

extern "C" int pnamp_set_form(pnamp_ctx *c, int form) {
  if (!c) return PNAMP_E_ARG;
  if (form < PNAMP_FORM_PLAIN || form > PNAMP_FORM_MIXED_BF16) return fail(c, PNAMP_E_ARG, "unknown form");
  c->form = form;
  return PNAMP_OK;
}

extern "C" int pnamp_states_separable(pnamp_ctx *c) { return c ? (c->separable ? 1 : 0) : PNAMP_E_ARG; }

extern "C" int pnamp_state_overlaps(pnamp_ctx *c, double *det) {
  if (!c || !det) return PNAMP_E_ARG;
  if (!c->nstates) return fail(c, PNAMP_E_STATE, "no states set");
  HIPCHK(c, hipMemcpy(det, c->d_det, sizeof(double) * c->nstates, hipMemcpyDeviceToHost));
  return PNAMP_OK;
}

static hipError_t launch_project(pnamp_ctx *c, int npairs, const ProjParams &pr) {
  const size_t J2m = 2 * (size_t)c->jmax + 1;
  if (c->fold == PNAMP_FOLD_PLAIN && (size_t)(c->jmax + 1) * J2m * J2m <= 256 * (size_t)PROJ_MAXO) {
    const size_t lds = sizeof(double2) * (J2m * (c->nalp + c->ngam) + (size_t)c->nalp * c->ngam +
                                          (size_t)c->nalp * J2m + J2m * J2m);
    if (lds > 48 * 1024) {
      const hipError_t e = hipFuncSetAttribute((const void *)k_project_sep, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_project_sep, dim3(npairs, pr.nobs), dim3(256), lds, c->stream, pr);
    return hipGetLastError();
  }
  const size_t lds = sizeof(double2) * (2 * (size_t)c->jmax + 1) * (c->nalp + c->ngam);
  if (lds > 48 * 1024) {
    const hipError_t e = hipFuncSetAttribute((const void *)k_project, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_project, dim3(npairs, pr.nobs), dim3(256), lds, c->stream, pr);
  return hipGetLastError();
}

static int ensure_pairs(pnamp_ctx *c, int np) {
  if (np <= c->pair_cap) return 0;
  int r = 0;
  const size_t ko = (size_t)NOBS * (c->jmax + 1) * (2 * c->jmax + 1) * (2 * c->jmax + 1);
  if ((r = dalloc(c, &c->d_q0, np)) || (r = dalloc(c, &c->d_q1, np)) || (r = dalloc(c, &c->d_pz, np)) ||
      (r = dalloc(c, &c->d_pn, np)) ||
      (r = dalloc(c, &c->d_zres, (size_t)np * c->NE * NOBS)) || (r = dalloc(c, &c->d_kern, (size_t)np * ko)) ||
      (r = dalloc(c, &c->d_status, (size_t)np * c->NE)) || (r = dalloc(c, &c->d_pair_status, (size_t)np)))
    return r;
  c->pair_cap = np;
  return 0;
}


// Arguments of the plain-form kernels for the pairs q0/q1 (device), writing
// zres/status per (pair, Euler point).
static PlainArgs plain_args(const pnamp_ctx *c, int zval, int nval, const int *q0, const int *q1, double2 *zres,
                            int *status) {
  PlainArgs a{};
  a.n = c->n; a.h = c->h; a.nalp = c->nalp; a.nbet = c->nbet; a.ngam = c->ngam; a.NE = c->NE; a.nphi = c->nphi;
  a.npack = c->npack; a.zval = zval; a.nval = nval;
  a.cscale = c->cscale; a.vac_scale = c->vac_scale; a.snpfaf = c->snpfaf; a.e0 = c->e0;
  a.q0 = q0; a.q1 = q1; a.pz = nullptr; a.pn = nullptr;
  a.U = c->d_U; a.V = c->d_V; a.Z = c->d_Z; a.det = c->d_det;
  a.alp = c->d_alp; a.gam = c->d_gam; a.rtab = c->d_rtab; a.pk_a = c->d_pk_a; a.pk_b = c->d_pk_b;
  a.blk_off = c->d_blk_off; a.blk_sz = c->d_blk_sz; a.blk_base = c->d_blk_base; a.twom = c->d_twom;
  a.mtab = c->d_mtab; a.t1b = c->d_t1b; a.lofs = c->d_plofs; a.lme = c->d_plme; a.nl = c->pl_n;
  a.lbytes = c->pl_bytes;
  a.q2 = nullptr; a.orbp = nullptr; a.nob2 = 0;
  a.zres = zres; a.status = status; a.ze2 = nullptr; a.zrho = nullptr;
  return a;
}

static ProjParams proj_params(const pnamp_ctx *c, const double2 *zres, const int *status, double2 *kern,
                              int *pair_status) {
  return ProjParams{c->nalp, c->nbet, c->ngam, c->NE, c->jmax, c->fold, c->npj, NOBS, c->d_alp, c->d_walp,
                    c->d_wbet, c->d_gam, c->d_wgam, c->d_ptab, zres, status, kern, pair_status};
}

// The batched kernels; with samples (device, [npairs][NE][nphi][nphi][2]) the
// factored forms also write the norm and energy of every gauge pair, and run
// their launches directly instead of through the captured graph.
struct E2Out {
  double *ze2, *qred, *rho1b;  // host outputs (any may be NULL)
};

static int ensure_e2(pnamp_ctx *c, int np) {
  if (np <= c->e2_cap) return 0;
  const size_t J2m = 2 * (size_t)c->jmax + 1, per_io = (size_t)(c->jmax + 1) * J2m * J2m;
  int r = 0;
  if ((r = dalloc(c, &c->d_ze2, (size_t)np * c->NE * 10)) || (r = dalloc(c, &c->d_zrho, (size_t)np * c->NE * c->nob2)) ||
      (r = dalloc(c, &c->d_iproj, (size_t)np * 10 * per_io)) ||
      (r = dalloc(c, &c->d_qred, (size_t)np * 2 * 5 * (c->jmax + 1) * (J2m + 4) * J2m)) ||
      (r = dalloc(c, &c->d_rho1b, (size_t)np * c->nob2)))
    return r;
  c->e2_cap = np;
  return 0;
}

static int kernels_impl(pnamp_ctx *c, int npairs, const int32_t *q0, const int32_t *q1, int zval, int nval,
                        double *zres, double *kern, double2 *samples, const E2Out *e2 = nullptr,
                        const int32_t *zv = nullptr, const int32_t *nv = nullptr) {
  if (!c || npairs <= 0 || !q0 || !q1) return PNAMP_E_ARG;
  const bool per_pair = zv != nullptr;
  if (per_pair && !nv) return PNAMP_E_ARG;
  if (!c->nstates) return fail(c, PNAMP_E_STATE, "no states set");
  if (e2) {
    if (!c->d_q2) return fail(c, PNAMP_E_STATE, "E2 kernels need the Q2 matrices (pnamp_set_q2)");
    if (c->fold != PNAMP_FOLD_PLAIN)
      return fail(c, PNAMP_E_ARG, "E2 kernels are projected with the plain fold only");
    if (npairs > 65535) return fail(c, PNAMP_E_SIZE, "at most 65535 pairs per call");
    int r2 = ensure_e2(c, npairs);
    if (r2) return r2;
  }
  for (int i = 0; i < npairs; ++i)
    if (q0[i] < 0 || q1[i] < 0 || q0[i] >= c->nstates || q1[i] >= c->nstates)
      return fail(c, PNAMP_E_ARG, "pair index out of range");
  if (npairs > 65535) return fail(c, PNAMP_E_SIZE, "at most 65535 pairs per call");
  int r = ensure_pairs(c, npairs);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(c->d_q0, q0, sizeof(int) * npairs, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_q1, q1, sizeof(int) * npairs, hipMemcpyHostToDevice, c->stream));
  if (per_pair) {
    HIPCHK(c, hipMemcpyAsync(c->d_pz, zv, sizeof(int) * npairs, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_pn, nv, sizeof(int) * npairs, hipMemcpyHostToDevice, c->stream));
  }
  if (c->form != PNAMP_FORM_PLAIN) {
    if (!c->separable)
      return fail(c, PNAMP_E_ARG, "factored and mixed forms need states without proton-neutron mixing");
    if (!c->fac_list_ok) return fail(c, PNAMP_E_ARG, c->fac_reason);
    const int mix = (c->form == PNAMP_FORM_FACTORED) ? MIX_FP64 : (c->form == PNAMP_FORM_MIXED_FP32 ? MIX_FP32 : MIX_BF16);
    const int h = c->h, hh = h * h, kp = c->kp;
    if (c->nphi > NPHI_MAX) return fail(c, PNAMP_E_SIZE, "factored form supports nphi <= 16");
    const size_t comb_lds = sizeof(double2) * (size_t)c->nphi * hh + sizeof(int) * (size_t)c->pn_ncol;
    if (comb_lds > 150 * 1024) return fail(c, PNAMP_E_SIZE, "nphi h^2 complex exceeds the combine LDS");
    if (hh > ((FAC_HMAX * FAC_HMAX + NTC - 1) / NTC) * NTC)
      return fail(c, PNAMP_E_SIZE, "h^2 exceeds the combine rows per workgroup");
    if (h > FAC_HMAX || (h & 1)) return fail(c, PNAMP_E_SIZE, "factored form: species size h must be even and <= 22");
    // per pair: rho [NE][2][nphi][h^2], setup blocks [NE][2][5][h^2], per-phi scalars and status;
    // mixed forms add the GEMM operands (4 x 2 NE nphi columns x kp, 4 bytes per element),
    // A01 [NE][2][nphi][h^2] and G_pn [NE][nphi][h^2] (FP32 complex); chunks up to 24 GiB
    const size_t rho_pp = (size_t)c->NE * 2 * c->nphi * hh, sb_pp = (size_t)c->NE * 2 * 5 * hh;
    const size_t out_pp = (size_t)c->NE * 2 * c->nphi * FAC_OUT, st_pp = (size_t)c->NE * 2 * c->nphi;
    const size_t cols_pp = (size_t)2 * c->NE * c->nphi;  // GEMM columns per pair and species
    size_t bytes_pp = 16 * (rho_pp + sb_pp + out_pp) + 4 * st_pp;
    if (mix != MIX_FP64) bytes_pp += 4 * 4 * cols_pp * kp + 16 * rho_pp + 8 * (size_t)c->NE * c->nphi * hh;
    const size_t chunk_max = std::max<size_t>(1, ((size_t)24 << 30) / bytes_pp);
    const size_t chunk_t = mix == MIX_FP64 ? std::max<size_t>(1, chunk_max / (size_t)std::max(1, c->tune.chunk_div)) : chunk_max;
    const int chunk = (int)std::min<size_t>((size_t)npairs, chunk_t);
    if ((size_t)chunk > c->fac_cap) {
      if ((r = dalloc(c, &c->d_frho, rho_pp * chunk)) || (r = dalloc(c, &c->d_sbuf, sb_pp * chunk)) ||
          (r = dalloc(c, &c->d_fout, out_pp * chunk)) || (r = dalloc(c, &c->d_fstatus, st_pp * chunk)))
        return r;
      c->fac_cap = chunk;
      c->mix_cap = 0;
    }
    const size_t ncols_cap = (cols_pp * chunk + 127) / 128 * 128;
    if (mix != MIX_FP64 && (c->mix_cap < (size_t)chunk || c->mix_cap_mix != mix)) {
      // GEMM operands: rows k >= h^2 and columns past a chunk's last stay zero
      const size_t eb = (mix == MIX_FP32) ? 4 : 2;
      for (int q = 0; q < 2; ++q) {
        void **bufs[4] = {&c->d_brho[q], &c->d_brho_lo[q], &c->d_ba10[q], &c->d_ba10_lo[q]};
        for (int b = 0; b < 4; ++b) {
          free_void(c, bufs[b]);
          if (mix == MIX_FP32 && (b & 1)) continue;
          HIPCHK(c, hipMalloc(bufs[b], ncols_cap * kp * eb));
          HIPCHK(c, hipMemset(*bufs[b], 0, ncols_cap * kp * eb));
        }
      }
      if ((r = dalloc(c, &c->d_a01, rho_pp * chunk)) || (r = dalloc(c, &c->d_gpn, (size_t)2 * c->NE * c->nphi * hh * chunk)))
        return r;
      c->mix_cap = chunk;
      c->mix_cap_mix = mix;
    }
    if (mix != MIX_FP64 && c->va_mix != mix) {
      for (int q = 0; q < NVPROB; ++q)
        if ((r = upload_vmat(c, q, mix))) return r;
      c->va_mix = mix;
    }
    if (!combine_attr(c->nphi, comb_lds)) return fail(c, PNAMP_E_SIZE, "factored form supports nphi <= 16");
    const size_t setup_lds = sizeof(double2) * 3 * (size_t)hh;
    SetupParams sp;
    sp.n = c->n; sp.h = h; sp.nalp = c->nalp; sp.ngam = c->ngam; sp.npack = c->npack; sp.NE = c->NE;
    sp.alp = c->d_alp; sp.gam = c->d_gam; sp.Uf = c->d_Uf; sp.Vf = c->d_Vf; sp.Z = c->d_Z; sp.rtab = c->d_rtab;
    sp.pk_a = c->d_pk_a; sp.pk_b = c->d_pk_b; sp.blk_off = c->d_blk_off; sp.blk_sz = c->d_blk_sz;
    sp.blk_base = c->d_blk_base; sp.twom = c->d_twom; sp.sbuf = c->d_sbuf;
    PhiParams fp;
    fp.n = c->n; fp.nphi = c->nphi; fp.NE = c->NE; fp.cscale = c->cscale; fp.sbuf = c->d_sbuf;
    fp.UfT = c->d_UfT; fp.VfT = c->d_VfT; fp.Z = c->d_Z; fp.t1b = c->d_t1b; fp.mtab = c->d_mtab;
    fp.binv = c->d_binv; fp.pfb = c->d_pfb;
    const bool pfs = c->schur_allowed && c->pfs_any, both = pfs && !c->pfs_all && c->tune.pf_waves == 0;
    fp.pfs_ok = both ? c->d_pfs_ok : nullptr;
    fp.jtab = (const double2 *)c->d_jtab;
    for (int q = 0; q < 2; ++q) {
      fp.lofs[q] = c->d_lofs[q]; fp.nl[q] = c->nl[q];
    }
    fp.out = c->d_fout; fp.rho = c->d_frho; fp.status = c->d_fstatus;
    for (int q = 0; q < 2; ++q) {
      fp.brho[q] = c->d_brho[q]; fp.brho_lo[q] = c->d_brho_lo[q];
      fp.ba10[q] = c->d_ba10[q]; fp.ba10_lo[q] = c->d_ba10_lo[q];
    }
    fp.a01 = c->d_a01;
    VGemmParams vp;
    vp.h = h; vp.nphi = c->nphi; vp.out = c->d_fout;
    for (int q = 0; q < NVPROB; ++q) {
      VProblem &P = vp.prob[q];
      const int spc = (q == 4) ? 1 : (q & 1);
      P.ahi = c->d_va[q]; P.alo = c->d_va_lo[q];
      P.bhi = (q < 2 || q == 4) ? c->d_brho[spc] : c->d_ba10[spc];
      P.blo = (q < 2 || q == 4) ? c->d_brho_lo[spc] : c->d_ba10_lo[spc];
      P.dot = (q < 2) ? c->d_frho : c->d_a01;
      P.g = c->d_gpn;
      P.mode = (q == 4) ? 1 : 0;
      P.species = spc;
      P.slot = (q < 2) ? 1 : 8;
      P.ncols = 0;
    }
    CombParams cp;
    cp.h = h; cp.NE = c->NE; cp.zval = zval; cp.nval = nval;
    cp.vac_scale = c->vac_scale; cp.snpfaf = c->snpfaf; cp.sgn_h = (h & 1) ? -1.0 : 1.0; cp.e0 = c->e0;
    cp.det = c->d_det; cp.pn_rows = c->d_pn_rows; cp.pn_blk = c->d_pn_blk; cp.pn_ycol = c->d_pn_ycol;
    cp.pn_w = c->d_pn_w; cp.pn_ncol = c->pn_ncol;
    cp.out = c->d_fout; cp.rho = c->d_frho; cp.fstatus = c->d_fstatus; cp.gpn = (const float2 *)c->d_gpn;
    cp.samples = nullptr;
    if (c->fac_graph && (c->g_np != npairs || c->g_form != c->form || c->g_zn != (int)per_pair ||
                         (!per_pair && (c->g_z != zval || c->g_n != nval))))
      drop_graph(c);
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    const bool direct = samples != nullptr || e2 != nullptr;
    FacE2Params ep;
    ep.n = c->n; ep.h = h; ep.NE = c->NE; ep.nphi = c->nphi; ep.zval = zval; ep.nval = nval; ep.nob2 = c->nob2;
    ep.vac_scale = c->vac_scale; ep.snpfaf = c->snpfaf; ep.sgn_h = (h & 1) ? -1.0 : 1.0; ep.det = c->d_det;
    ep.out = c->d_fout; ep.rho = c->d_frho; ep.q2 = c->d_q2; ep.orbp = c->d_orbp;
    if (direct || !c->fac_graph) {
      // the launches of all chunks: captured once into a graph that later
      // calls of the same shape and form replay (the pair lists are copied
      // into the same device buffers before each replay), or issued directly
      // when gauge-point samples are requested
      if (!direct) HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (int p0 = 0; p0 < npairs; p0 += chunk) {
        const int np = std::min(chunk, npairs - p0);
        sp.q0 = c->d_q0 + p0; sp.q1 = c->d_q1 + p0;
        fp.q0 = c->d_q0 + p0; fp.q1 = c->d_q1 + p0;
        cp.q0 = c->d_q0 + p0; cp.q1 = c->d_q1 + p0;
        cp.pz = per_pair ? c->d_pz + p0 : nullptr;
        cp.pn = per_pair ? c->d_pn + p0 : nullptr;
        cp.zres = c->d_zres + (size_t)p0 * c->NE * NOBS;
        cp.status = c->d_status + (size_t)p0 * c->NE;
        if (direct) cp.samples = samples + (size_t)p0 * c->NE * c->nphi * c->nphi * 2;
        const size_t nwg = (size_t)np * c->NE * 2 * c->nphi;
        fp.nwg = (int)nwg;
        hipLaunchKernelGGL(k_fac_setup, dim3(c->NE, np, 2), dim3(NTS), setup_lds, c->stream, sp);
        launch_phi_h(h, mix, c->tune, pfs ? (both ? 2 : 1) : 0, (unsigned)((nwg + 7) / 8 * 8), c->stream, fp);
        if (mix != MIX_FP64) {
          const int ncols = (int)(cols_pp * np);
          for (int q = 0; q < NVPROB; ++q) vp.prob[q].ncols = ncols;
          launch_vgemm(h, mix, dim3((ncols + 127) / 128, NVPROB), c->stream, vp);
        }
        launch_combine_np(c->nphi, mix != MIX_FP64, dim3(c->NE, np), comb_lds, c->stream, cp);
        if (e2) {
          ep.q0 = c->d_q0 + p0; ep.q1 = c->d_q1 + p0;
          ep.pz = per_pair ? c->d_pz + p0 : nullptr;
          ep.pn = per_pair ? c->d_pn + p0 : nullptr;
          ep.ze2 = c->d_ze2 + (size_t)p0 * c->NE * 10;
          ep.zrho = c->d_zrho + (size_t)p0 * c->NE * c->nob2;
          hipLaunchKernelGGL(k_fac_e2, dim3(c->NE, np), dim3(256), 0, c->stream, ep);
        }
      }
    }
    if (!direct && !c->fac_graph) {
      hipGraph_t graph = nullptr;
      const hipError_t cap = hipStreamEndCapture(c->stream, &graph);
      if (cap != hipSuccess) return fail(c, PNAMP_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(cap));
      size_t nodes = 0;
      if (hipGraphGetNodes(graph, nullptr, &nodes) == hipSuccess) c->graph_nodes = (int)nodes;
      const hipError_t ins = hipGraphInstantiate(&c->fac_graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ins != hipSuccess) {
        c->fac_graph = nullptr;
        return fail(c, PNAMP_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ins));
      }
      ++c->graph_captures;
      c->g_np = npairs;
      c->g_z = zval;
      c->g_n = nval;
      c->g_form = c->form;
      c->g_zn = per_pair ? 1 : 0;
    }
    if (!direct) {
      HIPCHK(c, hipGraphLaunch(c->fac_graph, c->stream));
      ++c->graph_replays;
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  } else {
    PlainArgs pa = plain_args(c, zval, nval, c->d_q0, c->d_q1, c->d_zres, c->d_status);
    if (per_pair) {
      pa.pz = c->d_pz;
      pa.pn = c->d_pn;
    }
    if (e2) {
      pa.q2 = c->d_q2; pa.orbp = c->d_orbp; pa.nob2 = c->nob2; pa.ze2 = c->d_ze2; pa.zrho = c->d_zrho;
    }
    // units (pair, Euler point) in chunks whose work buffer stays under 4 GiB
    const size_t ub = plain_unit_bytes(pa), units = (size_t)npairs * c->NE;
    if (units > (size_t)INT32_MAX / 2) return fail(c, PNAMP_E_SIZE, "too many (pair, Euler point) units in one call");
    const size_t cu = std::max<size_t>(1, std::min(units, ((size_t)4 << 30) / ub));
    if (c->pwork_bytes < cu * ub) {
      free_void(c, &c->d_pwork);
      c->pwork_bytes = 0;
      HIPCHK(c, hipMalloc(&c->d_pwork, cu * ub));
      c->pwork_bytes = cu * ub;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    for (size_t u0 = 0; u0 < units; u0 += cu)
      HIPCHK(c, plain_run(pa, (int)u0, (int)std::min(cu, units - u0), c->d_pwork, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  }
  const ProjParams pr = proj_params(c, c->d_zres, c->d_status, c->d_kern, c->d_pair_status);
  HIPCHK(c, launch_project(c, npairs, pr));
  if (e2) {
    // the ten E2 observables through the plain projection, then the 3j sums
    ProjParams pe = proj_params(c, c->d_ze2, c->d_status, c->d_iproj, c->d_pair_status);
    pe.nobs = 10;
    HIPCHK(c, launch_project(c, npairs, pe));
    E2RedParams rp{c->jmax, c->NE, c->nob2, c->nalp, c->ngam, c->d_iproj, c->d_t3j, c->d_qred, c->d_zrho,
                   c->d_walp, c->d_wbet, c->d_wgam, c->d_orbscale, c->d_rho1b};
    hipLaunchKernelGGL(k_e2_reduce, dim3(npairs), dim3(256), 0, c->stream, rp);
    HIPCHK(c, hipGetLastError());
    const size_t J2m = 2 * (size_t)c->jmax + 1;
    if (e2->ze2)
      HIPCHK(c, hipMemcpyAsync(e2->ze2, c->d_ze2, sizeof(double2) * (size_t)npairs * c->NE * 10, hipMemcpyDeviceToHost,
                               c->stream));
    if (e2->qred)
      HIPCHK(c, hipMemcpyAsync(e2->qred, c->d_qred, sizeof(double2) * (size_t)npairs * 10 * (c->jmax + 1) * (J2m + 4) * J2m,
                               hipMemcpyDeviceToHost, c->stream));
    if (e2->rho1b)
      HIPCHK(c, hipMemcpyAsync(e2->rho1b, c->d_rho1b, sizeof(double2) * (size_t)npairs * c->nob2, hipMemcpyDeviceToHost,
                               c->stream));
  }
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  const size_t kper = (size_t)NOBS * (c->jmax + 1) * (2 * c->jmax + 1) * (2 * c->jmax + 1);
  if (kern)  // k_project writes every slot of the dense layout, zeros where |K| > J
    HIPCHK(c, hipMemcpyAsync(kern, c->d_kern, sizeof(double2) * kper * npairs, hipMemcpyDeviceToHost, c->stream));
  if (zres)
    HIPCHK(c, hipMemcpyAsync(zres, c->d_zres, sizeof(double2) * (size_t)npairs * c->NE * NOBS,
                             hipMemcpyDeviceToHost, c->stream));
  std::vector<int> pst(npairs);
  HIPCHK(c, hipMemcpyAsync(pst.data(), c->d_pair_status, sizeof(int) * npairs, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipEventElapsedTime(&c->ms_points, c->ev[0], c->ev[1]));
  HIPCHK(c, hipEventElapsedTime(&c->ms_total, c->ev[0], c->ev[3]));
  for (int i = 0; i < npairs; ++i)
    if (pst[i]) return fail(c, pst[i], "singular T (ROKAPPAphiomega) in pair " + std::to_string(i));
  return PNAMP_OK;
}

extern "C" int pnamp_kernels_batched(pnamp_ctx *c, int npairs, const int32_t *q0, const int32_t *q1, int zval,
                                     int nval, double *zres, double *kern) {
  return kernels_impl(c, npairs, q0, q1, zval, nval, zres, kern, nullptr);
}

// (j1 j2 j3; m1 m2 m3), arguments doubled, Racah formula
static double threej_host(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3) {
  if (tm1 + tm2 + tm3 != 0 || std::abs(tm1) > tj1 || std::abs(tm2) > tj2 || std::abs(tm3) > tj3) return 0.0;
  if (((tj1 + tm1) & 1) || ((tj2 + tm2) & 1) || ((tj3 + tm3) & 1)) return 0.0;
  if (tj3 > tj1 + tj2 || tj3 < std::abs(tj1 - tj2) || ((tj1 + tj2 + tj3) & 1)) return 0.0;
  auto f = [](int k) { return std::tgamma(k + 1.0); };
  const int a = (tj1 + tj2 - tj3) / 2, b = (tj1 - tj2 + tj3) / 2, cc = (-tj1 + tj2 + tj3) / 2;
  const double pre = std::sqrt(f(a) * f(b) * f(cc) / f((tj1 + tj2 + tj3) / 2 + 1) * f((tj1 + tm1) / 2) *
                               f((tj1 - tm1) / 2) * f((tj2 + tm2) / 2) * f((tj2 - tm2) / 2) * f((tj3 + tm3) / 2) *
                               f((tj3 - tm3) / 2));
  const int t1 = (tj3 - tj2 + tm1) / 2, t2 = (tj3 - tj1 - tm2) / 2;
  const int kmin = std::max({0, -t1, -t2}), kmax = std::min({a, (tj1 - tm1) / 2, (tj2 + tm2) / 2});
  double sum = 0.0;
  for (int k = kmin; k <= kmax; ++k) {
    const double den = f(k) * f(a - k) * f((tj1 - tm1) / 2 - k) * f((tj2 + tm2) / 2 - k) * f(t1 + k) * f(t2 + k);
    sum += (k & 1) ? -1.0 / den : 1.0 / den;
  }
  return (((tj1 - tj2 - tm3) / 2) & 1 ? -1.0 : 1.0) * pre * sum;
}

extern "C" int pnamp_set_q2(pnamp_ctx *c, const double *q2) {
  if (!c || !q2) return PNAMP_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const int n = c->n, h = c->h, jm = c->jmax;
  // orbits of one species half: contiguous runs m = j .. -j
  std::vector<int> first;
  for (int i = 0; i < h; i += c->twoj_h[i] + 1) first.push_back(i);
  const int norb = (int)first.size();
  if (2 * norb * norb > NOB2_MAX) return fail(c, PNAMP_E_SIZE, "more than 8 orbits per species");
  std::vector<int4> op;
  std::vector<double> osc;
  const double w0 = (c->nalp * c->ngam == 1) ? 0.5 : 1.0 / (8.0 * M_PI * M_PI);
  for (int sp = 0; sp < 2; ++sp)
    for (int a = 0; a < norb; ++a)
      for (int b = 0; b < norb; ++b) {
        const int fa = first[a], fb = first[b];
        const bool same = c->twoj_h[fa] == c->twoj_h[fb] && c->l_h[fa] == c->l_h[fb];
        op.push_back(make_int4(sp, sp * h + fa, sp * h + fb, same ? c->twoj_h[fa] + 1 : 0));
        osc.push_back(w0 / std::sqrt(c->twoj_h[fa] + 1.0));
      }
  // T(J1 - J2, J2, K1, mu) = (2 J1 + 1) (-1)^(J1 - K1) (J1 2 J2; -K1 mu K1 - mu)
  const int K1d = 2 * jm + 5;
  std::vector<double> t3((size_t)5 * (jm + 1) * K1d * 5, 0.0);
  for (int dJ = -2; dJ <= 2; ++dJ)
    for (int J2 = 0; J2 <= jm; ++J2) {
      const int J1 = J2 + dJ;
      if (J1 < 0) continue;
      for (int K1 = -J1; K1 <= J1; ++K1)
        for (int mu = -2; mu <= 2; ++mu)
          t3[(((size_t)(dJ + 2) * (jm + 1) + J2) * K1d + (K1 + jm + 2)) * 5 + (mu + 2)] =
              (2.0 * J1 + 1.0) * (((J1 - K1) & 1) ? -1.0 : 1.0) *
              threej_host(2 * J1, 4, 2 * J2, -2 * K1, 2 * mu, 2 * (K1 - mu));
    }
  int r = 0;
  if ((r = dupload(c, &c->d_q2, q2, (size_t)5 * n * n)) || (r = dupload(c, &c->d_orbp, op.data(), op.size())) ||
      (r = dupload(c, &c->d_orbscale, osc.data(), osc.size())) || (r = dupload(c, &c->d_t3j, t3.data(), t3.size())))
    return r;
  c->norb = norb;
  c->nob2 = (int)op.size();
  c->e2_cap = 0;
  return PNAMP_OK;
}

extern "C" int pnamp_norb(pnamp_ctx *c) {
  if (!c) return PNAMP_E_ARG;
  int norb = 0;
  for (int i = 0; i < c->h; i += c->twoj_h[i] + 1) ++norb;
  return norb;
}

extern "C" int pnamp_kernels_e2(pnamp_ctx *c, int npairs, const int32_t *q0, const int32_t *q1, int zval, int nval,
                                double *zres, double *kern, double *ze2, double *qred, double *rho1b) {
  const E2Out e2{ze2, qred, rho1b};
  return kernels_impl(c, npairs, q0, q1, zval, nval, zres, kern, nullptr, &e2);
}

extern "C" int pnamp_kernels_batched_zn(pnamp_ctx *c, int npairs, const int32_t *q0, const int32_t *q1,
                                        const int32_t *zval, const int32_t *nval, double *zres, double *kern) {
  if (!zval || !nval) return PNAMP_E_ARG;
  return kernels_impl(c, npairs, q0, q1, 0, 0, zres, kern, nullptr, nullptr, zval, nval);
}

extern "C" int pnamp_mesh_samples(pnamp_ctx *c, int npairs, const int32_t *q0, const int32_t *q1, int zval, int nval,
                                  double *samples) {
  if (!c || !samples || npairs <= 0) return PNAMP_E_ARG;
  if (c->form == PNAMP_FORM_PLAIN) return fail(c, PNAMP_E_ARG, "mesh samples need a factored or mixed form");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t cnt = (size_t)npairs * c->NE * c->nphi * c->nphi * 2;
  double2 *d = nullptr;
  HIPCHK(c, hipMalloc((void **)&d, cnt * sizeof(double2)));
  const int r = kernels_impl(c, npairs, q0, q1, zval, nval, nullptr, nullptr, d);
  hipError_t e = hipSuccess;
  if (r == PNAMP_OK) e = hipMemcpy(samples, d, cnt * sizeof(double2), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, PNAMP_E_DEVICE, std::string("samples copy: ") + hipGetErrorString(e));
  return r;
}

extern "C" int pnamp_project_zres(pnamp_ctx *c, int npairs, const double *zres, double *kern) {
  if (!c || npairs <= 0 || !zres || !kern) return PNAMP_E_ARG;
  if (npairs > 65535) return fail(c, PNAMP_E_SIZE, "at most 65535 pairs per call");
  int r = ensure_pairs(c, npairs);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(c->d_zres, zres, sizeof(double2) * (size_t)npairs * c->NE * NOBS, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_status, 0, sizeof(int) * (size_t)npairs * c->NE, c->stream));
  const ProjParams pr = proj_params(c, c->d_zres, c->d_status, c->d_kern, c->d_pair_status);
  HIPCHK(c, launch_project(c, npairs, pr));
  const size_t kper = (size_t)NOBS * (c->jmax + 1) * (2 * c->jmax + 1) * (2 * c->jmax + 1);
  HIPCHK(c, hipMemcpyAsync(kern, c->d_kern, sizeof(double2) * kper * npairs, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PNAMP_OK;
}

extern "C" int pnamp_kernel_pair(pnamp_ctx *c, int q0, int q1, int zval, int nval, double *zres, double *kern) {
  const int32_t a = q0, b = q1;
  return pnamp_kernels_batched(c, 1, &a, &b, zval, nval, zres, kern);
}

extern "C" int pnamp_kernels_batched_noalloc(pnamp_ctx *c, int npairs, const int32_t *q0, const int32_t *q1,
                                             int zval, int nval, double *zres, double *kern) {
  if (!c || npairs <= 0 || !q0 || !q1) return PNAMP_E_ARG;
  if (!c->nstates) return fail(c, PNAMP_E_STATE, "no states set");
  if (!c->scratch) return fail(c, PNAMP_E_STATE, "no scratch address for this device");
  for (int i = 0; i < npairs; ++i)
    if (q0[i] < 0 || q1[i] < 0 || q0[i] >= c->nstates || q1[i] >= c->nstates)
      return fail(c, PNAMP_E_ARG, "pair index out of range");
  // scratch: the first half for the pair chunk (q0, q1, ZRES, per-Euler
  // status, projected kernels, per-pair status, each 256-byte aligned), the
  // second half for the plain kernels' work buffer (units in sub-chunks)
  const size_t kper = (size_t)NOBS * (c->jmax + 1) * (2 * c->jmax + 1) * (2 * c->jmax + 1);
  const size_t per_pair = 2 * sizeof(int) + sizeof(double2) * (size_t)c->NE * NOBS + sizeof(int) * (size_t)c->NE +
                          sizeof(double2) * kper + sizeof(int);
  const size_t half = (size_t)PNAMP_NOALLOC_SCRATCH_BYTES / 2, cap = half - 6 * 256;
  if (per_pair > cap)
    return fail(c, PNAMP_E_SIZE, "one pair needs " + std::to_string(per_pair) + " bytes of scratch; the static array holds " +
                                     std::to_string((size_t)PNAMP_NOALLOC_SCRATCH_BYTES) + " (half for the pairs)");
  const int chunk = (int)std::min<size_t>({cap / per_pair, (size_t)npairs, (size_t)65535});
  HIPCHK(c, hipSetDevice(c->device));
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  char *base = (char *)c->scratch;
  int *sq0 = (int *)base;
  int *sq1 = (int *)(base + al(sizeof(int) * chunk));
  double2 *szr = (double2 *)((char *)sq1 + al(sizeof(int) * chunk));
  int *sst = (int *)((char *)szr + al(sizeof(double2) * (size_t)chunk * c->NE * NOBS));
  double2 *sk = (double2 *)((char *)sst + al(sizeof(int) * (size_t)chunk * c->NE));
  int *spst = (int *)((char *)sk + al(sizeof(double2) * kper * chunk));
  void *swork = base + half;
  PlainArgs pa = plain_args(c, zval, nval, sq0, sq1, szr, sst);
  const size_t ub = plain_unit_bytes(pa);
  if (ub > half)
    return fail(c, PNAMP_E_SIZE, "one (pair, Euler point) unit needs " + std::to_string(ub) + " bytes of work buffer; " +
                                     std::to_string(half) + " available");
  const size_t wu = std::max<size_t>(1, half / ub);  // units per launch
  std::vector<int> pst(chunk);
  for (int p0 = 0; p0 < npairs; p0 += chunk) {
    const int np = std::min(chunk, npairs - p0);
    HIPCHK(c, hipMemcpyAsync(sq0, q0 + p0, sizeof(int) * np, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(sq1, q1 + p0, sizeof(int) * np, hipMemcpyHostToDevice, c->stream));
    // k_plain_sum writes every ZRES and status entry of the chunk and
    // k_project every kernel and pair-status entry; each k_plain_gauge
    // workgroup writes its own gauge-point slots before k_plain_sum reads
    // them: nothing left by an earlier call is read
    const size_t units = (size_t)np * c->NE;
    for (size_t u0 = 0; u0 < units; u0 += wu)
      HIPCHK(c, plain_run(pa, (int)u0, (int)std::min(wu, units - u0), swork, c->stream));
    HIPCHK(c, launch_project(c, np, proj_params(c, szr, sst, sk, spst)));
    if (kern)
      HIPCHK(c, hipMemcpyAsync(kern + 2 * kper * p0, sk, sizeof(double2) * kper * np, hipMemcpyDeviceToHost, c->stream));
    if (zres)
      HIPCHK(c, hipMemcpyAsync(zres + 2 * (size_t)c->NE * NOBS * p0, szr, sizeof(double2) * (size_t)np * c->NE * NOBS,
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pst.data(), spst, sizeof(int) * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < np; ++i)
      if (pst[i]) return fail(c, pst[i], "singular T (ROKAPPAphiomega) in pair " + std::to_string(p0 + i));
  }
  return PNAMP_OK;
}

// Device buffers of one Hill-Wheeler solve, freed on scope exit.
struct HwBufs {
  std::vector<void *> p;
  ~HwBufs() {
    for (void *x : p)
      if (x) (void)hipFree(x);
  }
  template <class T>
  bool alloc(T **q, size_t cnt) {
    const hipError_t e = hipMalloc((void **)q, (cnt ? cnt : 1) * sizeof(T));
    p.push_back(*q);
    return e == hipSuccess;
  }
};

// The first half of HNDIAG (Codes/GCM/GCM_HNDiag.f90:25-77): diagonalise -N
// (EN descending), keep the M leading natural states with EN(i)/EN(1) >= eps
// (the first failure ends the count), at most nos of them, R = D / sqrt(EN)
// (dim x M), G = R^T H R (M x M).  Leaves N, H, R, G on the device.
static int hw_prepare(pnamp_ctx *c, HwBufs &B, int dim, const double *N, const double *H, double eps, int nos,
                      double *EN, int *Mout, double **dN, double **dH, double **dR, double **dG) {
  const size_t dd = (size_t)dim * dim;
  double *dA = nullptr, *dV = nullptr, *dD = nullptr, *dw = nullptr;
  int *dsw = nullptr;
  if (!B.alloc(&dA, dd) || !B.alloc(&dV, dd) || !B.alloc(dN, dd) || !B.alloc(dH, dd) || !B.alloc(&dsw, 1))
    return fail(c, PNAMP_E_DEVICE, "hipMalloc failed in the Hill-Wheeler solve");
  std::vector<double> negN(dd);
  for (size_t i = 0; i < dd; ++i) negN[i] = -N[i];
  HIPCHK(c, hipMemcpyAsync(dA, negN.data(), sizeof(double) * dd, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(*dN, N, sizeof(double) * dd, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(*dH, H, sizeof(double) * dd, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(1024), 0, c->stream, dA, dV, dim, dsw);
  HIPCHK(c, hipGetLastError());
  std::vector<double> A(dd), Vv(dd);
  HIPCHK(c, hipMemcpyAsync(A.data(), dA, sizeof(double) * dd, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Vv.data(), dV, sizeof(double) * dd, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // ascending eigenvalues of -N == descending EN
  std::vector<int> ord(dim);
  for (int i = 0; i < dim; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return A[x + (size_t)x * dim] < A[y + (size_t)y * dim]; });
  for (int i = 0; i < dim; ++i) EN[i] = -A[ord[i] + (size_t)ord[i] * dim];
  int M = 0;
  while (M < dim && EN[M] / EN[0] >= eps) ++M;
  if (nos > 0 && nos < M) M = nos;
  *Mout = M;
  if (M <= 0) return PNAMP_OK;
  std::vector<double> D((size_t)dim * M), w(M);
  for (int k = 0; k < M; ++k) {
    w[k] = EN[k];
    for (int i = 0; i < dim; ++i) D[i + (size_t)k * dim] = Vv[i + (size_t)ord[k] * dim];
  }
  if (!B.alloc(&dD, (size_t)dim * M) || !B.alloc(&dw, M) || !B.alloc(dR, (size_t)dim * M) || !B.alloc(dG, (size_t)M * M))
    return fail(c, PNAMP_E_DEVICE, "hipMalloc failed in the Hill-Wheeler solve");
  HIPCHK(c, hipMemcpyAsync(dD, D.data(), sizeof(double) * D.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dw, w.data(), sizeof(double) * M, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_scale_cols, dim3((dim * M + 255) / 256), dim3(256), 0, c->stream, dD, dw, *dR, dim, M);
  hipLaunchKernelGGL(k_congruence, dim3((M * M + 255) / 256), dim3(256), 0, c->stream, *dH, *dR, *dG, dim, M);
  HIPCHK(c, hipGetLastError());
  return PNAMP_OK;
}

extern "C" int pnamp_hw_solve(pnamp_ctx *c, int dim, const double *N, const double *H, double eps, int nos,
                              double *EN, double *E, double *F, int *nkept) {
  if (!c || dim <= 0 || !N || !H || !EN || !E || !nkept) return PNAMP_E_ARG;
  if (dim > 1022) return fail(c, PNAMP_E_SIZE, "HW dimension above 1022");
  HIPCHK(c, hipSetDevice(c->device));
  HwBufs B;
  double *dN = nullptr, *dH = nullptr, *dR = nullptr, *dG = nullptr, *dV = nullptr, *dF = nullptr;
  int *dsw = nullptr;
  int M = 0;
  const int r = hw_prepare(c, B, dim, N, H, eps, nos, EN, &M, &dN, &dH, &dR, &dG);
  if (r) return r;
  *nkept = M;
  if (M <= 0) return 1;
  if (!B.alloc(&dV, (size_t)M * M) || !B.alloc(&dsw, 1) || (F && !B.alloc(&dF, (size_t)dim * M)))
    return fail(c, PNAMP_E_DEVICE, "hipMalloc failed in pnamp_hw_solve");
  hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(1024), 0, c->stream, dG, dV, M, dsw);
  HIPCHK(c, hipGetLastError());
  std::vector<double> G((size_t)M * M), gv((size_t)M * M);
  HIPCHK(c, hipMemcpyAsync(G.data(), dG, sizeof(double) * G.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(gv.data(), dV, sizeof(double) * gv.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int> o2(M);
  for (int i = 0; i < M; ++i) o2[i] = i;
  std::stable_sort(o2.begin(), o2.end(), [&](int x, int y) { return G[x + (size_t)x * M] < G[y + (size_t)y * M]; });
  for (int k = 0; k < M; ++k) E[k] = G[o2[k] + (size_t)o2[k] * M];
  if (F) {
    // FF = RR GG (MAB in HNDIAG), eigenvectors in ascending order of E
    std::vector<double> gs((size_t)M * M);
    for (int k = 0; k < M; ++k)
      for (int l = 0; l < M; ++l) gs[l + (size_t)k * M] = gv[l + (size_t)o2[k] * M];
    HIPCHK(c, hipMemcpyAsync(dV, gs.data(), sizeof(double) * gs.size(), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_matmul, dim3((dim * M + 255) / 256), dim3(256), 0, c->stream, dR, dV, dF, dim, M, M);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(F, dF, sizeof(double) * (size_t)dim * M, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return PNAMP_OK;
}

extern "C" int pnamp_hw_presolve(pnamp_ctx *c, int dim, const double *N, const double *H, double eps, int kmax,
                                 double *EN, int *m, double *E, double *var, int *nos_best, double *e_best) {
  if (!c || dim <= 0 || !N || !H || !EN || !m || !E || !var || !nos_best || !e_best || kmax <= 0)
    return PNAMP_E_ARG;
  if (dim > 1022) return fail(c, PNAMP_E_SIZE, "HW dimension above 1022");
  HIPCHK(c, hipSetDevice(c->device));
  HwBufs B;
  double *dN = nullptr, *dH = nullptr, *dR = nullptr, *dG = nullptr, *dwork = nullptr, *dE = nullptr, *dvar = nullptr;
  int M = 0;
  const int r = hw_prepare(c, B, dim, N, H, eps, 0, EN, &M, &dN, &dH, &dR, &dG);
  if (r) return r;
  *m = M;
  for (int k = 0; k < kmax; ++k) {
    nos_best[k] = 0;
    e_best[k] = 0.0;
  }
  if (M <= 0) return 1;
  const size_t wsz = 2 * ((size_t)M * (M + 1) * (2 * M + 1) / 6);
  if (!B.alloc(&dwork, wsz) || !B.alloc(&dE, (size_t)M * kmax) || !B.alloc(&dvar, (size_t)M * kmax))
    return fail(c, PNAMP_E_DEVICE, "hipMalloc failed in pnamp_hw_presolve");
  NosParams np{dim, M, kmax, dN, dH, dR, dG, dwork, dE, dvar};
  hipLaunchKernelGGL(k_nos_scan, dim3(M), dim3(256), 0, c->stream, np);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(E, dE, sizeof(double) * (size_t)M * kmax, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(var, dvar, sizeof(double) * (size_t)M * kmax, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // PreSolve_HWG's choice (Codes/GCM/GCM_HWG.f90:60-79): per level the NOS of
  // least variance, starting from the reference's initial value 100 and
  // updating on a strictly smaller variance; NOS > M repeat NOS = M
  for (int k = 0; k < kmax; ++k) {
    double best = 100.0;
    for (int b = k; b < M; ++b) {
      const double q2 = var[(size_t)b * kmax + k];
      if (std::fabs(best) > std::fabs(q2)) {
        best = q2;
        nos_best[k] = b + 1;
        e_best[k] = E[(size_t)b * kmax + k];
      }
    }
  }
  return PNAMP_OK;
}

extern "C" int pnamp_last_timing(pnamp_ctx *c, float *ms_points, float *ms_total) {
  if (!c) return PNAMP_E_ARG;
  if (ms_points) *ms_points = c->ms_points;
  if (ms_total) *ms_total = c->ms_total;
  return PNAMP_OK;
}

extern "C" int pnamp_autotune(pnamp_ctx *c, int npairs, const int32_t *q0, const int32_t *q1, int zval, int nval,
                              int *choice, int *cand, float *ms, int *ncand) {
  if (!c || npairs <= 0 || !q0 || !q1) return PNAMP_E_ARG;
  if (c->form != PNAMP_FORM_FACTORED) return fail(c, PNAMP_E_ARG, "autotune: the FP64 factored form only");
  struct Cand {
    int pf, rest, div;
  };
  std::vector<Cand> cands;
  const int h = c->h;
  // Pfaffian: the Schur route (0) where the states allow it, the 2h x 2h
  // kernel at 3, 4, 5 waves per SIMD (h >= 8)
  const bool schur = c->schur_allowed && c->pfs_any;
  std::vector<int> pfs;
  if (schur) pfs.push_back(0);
  if (h >= 8) pfs.insert(pfs.end(), {3, 4, 5});
  for (int pf : pfs)
    for (int rest : {0, 4}) {
      if (rest == 4 && (h > 16 || h < 8)) continue;
      FacTune t;
      t.pf_waves = pf;
      t.rest_waves = rest;
      if (phi_scratch_h(h, t, schur) != 0) continue;  // a variant that spills registers is no candidate
      for (int div : {1, 2, 4}) cands.push_back({pf, rest, div});
    }
  if (cands.empty()) cands.push_back({0, 0, 1});
  const int pf_dflt = schur ? 0 : pf_default_waves(h);
  if ((int)cands.size() > PNAMP_TUNE_MAX) return fail(c, PNAMP_E_SIZE, "autotune: too many candidates");
  const FacTune keep = c->tune;
  float best = 0.0f;
  int ib = -1;
  for (size_t i = 0; i < cands.size(); ++i) {
    c->tune.pf_waves = cands[i].pf;
    c->tune.rest_waves = cands[i].rest;
    c->tune.chunk_div = cands[i].div;
    drop_graph(c);
    int r = kernels_impl(c, npairs, q0, q1, zval, nval, nullptr, nullptr, nullptr);  // captures
    if (!r) r = kernels_impl(c, npairs, q0, q1, zval, nval, nullptr, nullptr, nullptr);  // replays
    if (r) {
      c->tune = keep;
      drop_graph(c);
      return r;
    }
    if (ms) ms[i] = c->ms_points;
    if (cand) {
      cand[3 * i] = cands[i].pf ? cands[i].pf : pf_dflt;
      cand[3 * i + 1] = cands[i].rest ? cands[i].rest : (3 * h <= 64 ? 3 : 2);
      cand[3 * i + 2] = cands[i].div;
    }
    if (ib < 0 || c->ms_points < best) {
      best = c->ms_points;
      ib = (int)i;
    }
  }
  c->tune.pf_waves = cands[ib].pf;
  c->tune.rest_waves = cands[ib].rest;
  c->tune.chunk_div = cands[ib].div;
  drop_graph(c);
  if (choice) {
    choice[0] = cands[ib].pf ? cands[ib].pf : pf_dflt;
    choice[1] = cands[ib].rest ? cands[ib].rest : (3 * h <= 64 ? 3 : 2);  // the default
    choice[2] = cands[ib].div;
  }
  if (ncand) *ncand = (int)cands.size();
  return PNAMP_OK;
}

extern "C" int pnamp_set_schur_pfaffian(pnamp_ctx *c, int enable) {
  if (!c) return PNAMP_E_ARG;
  if (c->schur_allowed != (enable != 0)) drop_graph(c);
  c->schur_allowed = enable != 0;
  return PNAMP_OK;
}

extern "C" int pnamp_schur_pfaffian_active(pnamp_ctx *c) {
  return c ? ((c->pfs_any && c->schur_allowed) ? (c->pfs_all ? 2 : 1) : 0) : PNAMP_E_ARG;
}

extern "C" int pnamp_set_tune(pnamp_ctx *c, int pf_waves, int rest_waves, int chunk_div) {
  if (!c) return PNAMP_E_ARG;
  if ((pf_waves != 0 && pf_waves != 3 && pf_waves != 4 && pf_waves != 5) ||
      (rest_waves != 0 && rest_waves != 2 && rest_waves != 3 && rest_waves != 4) ||
      (chunk_div != 1 && chunk_div != 2 && chunk_div != 4))
    return fail(c, PNAMP_E_ARG, "pnamp_set_tune: waves 3, 4, 5 (Pfaffian), 2, 3, 4 (densities), chunk divisor 1, 2, 4");
  c->tune.pf_waves = pf_waves;                                      // 0: the default for h
  c->tune.rest_waves = (rest_waves == 4 && c->h <= 16) ? 4 : 0;     // 2 and 3 name the default
  c->tune.chunk_div = chunk_div;
  drop_graph(c);
  return PNAMP_OK;
}

extern "C" int pnamp_graph_info(pnamp_ctx *c, int *captures, int *replays, int *nodes) {
  if (!c) return PNAMP_E_ARG;
  if (captures) *captures = c->graph_captures;
  if (replays) *replays = c->graph_replays;
  if (nodes) *nodes = c->graph_nodes;
  return PNAMP_OK;
}

extern "C" int pnamp_n_euler(pnamp_ctx *c) { return c ? c->NE : PNAMP_E_ARG; }

extern "C" const char *pnamp_last_error(pnamp_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" void pnamp_destroy(pnamp_ctx *c) {
  if (!c) return;
  drop_graph(c);
  void *bufs[] = {c->d_alp, c->d_walp, c->d_bet, c->d_wbet, c->d_gam, c->d_wgam, c->d_rtab, c->d_ptab,
                  c->d_pk_a, c->d_pk_b, c->d_blk_off, c->d_blk_sz, c->d_blk_base, c->d_twoj, c->d_twom,
                  c->d_plofs, c->d_plme, c->d_pwork, c->d_t1b, c->d_jtab, c->d_U, c->d_V, c->d_Z, c->d_det, c->d_state_status, c->d_q0,
                  c->d_q1, c->d_pz, c->d_pn, c->d_zres, c->d_kern, c->d_status, c->d_pair_status,
                  c->d_lofs[0], c->d_lofs[1], c->d_Uf, c->d_Vf, c->d_binv, c->d_pfb, c->d_pfs_ok,
                  c->d_UfT, c->d_VfT, c->d_mtab, c->d_fout, c->d_frho,
                  c->d_sbuf, c->d_fstatus, c->d_pn_rows,
                  c->d_pn_blk, c->d_pn_ycol, c->d_pn_w};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (int q = 0; q < NVPROB; ++q) {
    if (c->d_va[q]) (void)hipFree(c->d_va[q]);
    if (c->d_va_lo[q]) (void)hipFree(c->d_va_lo[q]);
  }
  for (int q = 0; q < 2; ++q) {
    void *mb[4] = {c->d_brho[q], c->d_brho_lo[q], c->d_ba10[q], c->d_ba10_lo[q]};
    for (void *b : mb)
      if (b) (void)hipFree(b);
  }
  void *e2b[] = {c->d_q2, c->d_t3j, c->d_orbscale, c->d_orbp, c->d_ze2, c->d_zrho, c->d_iproj, c->d_qred, c->d_rho1b};
  for (void *b : e2b)
    if (b) (void)hipFree(b);
  if (c->d_a01) (void)hipFree(c->d_a01);
  if (c->d_gpn) (void)hipFree(c->d_gpn);
  for (auto e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}
