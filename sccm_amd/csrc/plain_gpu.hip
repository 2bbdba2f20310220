// plain_gpu.hip -- the plain-form mesh-point kernels of the PNAMP path for
// MI355X (gfx950): states of any kind (proton-neutron mixing included), FP64.
//
// One mesh point = (pair, Euler point, gauge angles phi_p, phi_n).  With
// A = diag(e^{i phi_p} R, e^{i phi_n} R) (UVphiomega, Codes/PNAMP/PNAMP_Proj.f90:396-455)
// and D = diag(e^{i phi_s(a)}) every gauge-dependent matrix is a phase
// rescaling of a gauge-independent one:
//   A Z1 A^T          = D M0 D,              M0 = R~ Z1 R~^T
//   T = V0^T A V1 + U0^T A* U1 = sum_s e^{i phi_s} X_s + e^{-i phi_s} Y_s,
//                     X_s = V0^T P_s R~ V1,  Y_s = U0^T P_s R~* U1
//   A V1 = D (R~ V1), A* U1 = D* (R~* U1)
// (R~ = diag(R, R), P_s the projector on species s).  k_plain_setup forms
// M0, X_p, X_n, Y_p, Y_n, R~ V1, R~* U1 once per (pair, Euler point);
// k_plain_gauge then needs only O(n^2) phase work per gauge point before the
// Pfaffian and the inverse.
//
// k_plain_gauge, one workgroup of NTH = 64 ceil(3N/32) threads per gauge
// point (N the size class of n):
//   * Pfaffian of the 2n x 2n skew matrix by ZPfaffianF's algorithm
//     (Codes/PNAMP/PNAMP_Proj.f90:218-350): full pivoting over the trailing
//     upper triangle (largest |S(i,j)|, first in row-major order on ties),
//     the two exchanges (NR I1), (NC I2), the rank-2 Schur update, Pf = 0 at
//     a pivot of modulus <= 1e-13.  The strict upper triangle lives in
//     registers, entry e in slot e / NTH of thread e % NTH, entries numbered
//     from the bottom-right corner so the trailing matrix is always the first
//     m(m-1)/2 entries.  The exchanges are not carried out: the trailing
//     entries read the old rows NR, NC, I1, I2 (four vectors in LDS) through
//     the composed transposition.  Two barriers per step.
//   * T^{-1}: W = [A V1; A* U1] T^{-1} by Gauss-Jordan column operations on the
//     stacked 3n x n matrix [T; A V1; A* U1] (pivot: largest |re|+|im| in the
//     current row of T among unused columns, as IZAMAX; the solution columns
//     come out permuted, the permutation is kept per register column).  Two
//     threads per stacked row hold one half of its columns each in
//     registers; one barrier per step (row k of T goes through LDS).
//   * rho = W1 V0^T, kappa10 = W1 U0^T, kappa01 = -W2 V0^T
//     (ROKAPPAphiomega, PNAMP_Proj.f90:459-551) straight from the registers;
//     rho to LDS, kappa only as A01 = kappa01 - kappa01^T (strict upper) and
//     A10 / 4 = (kappa10 - kappa10^T) / 4 (strict lower) of one n x n LDS
//     block, each entry the sum of exactly two LDS atomic adds onto zero
//     (so the result does not depend on their order).
//   * the energy (HFB_Energy_COMPLEX, PNAMP_Energy_OpenMP.f90:5-100) from the
//     regrouped list: (1/4) V [r_ijkl - r_jikl - r_ijlk + r_jilk + d (...)] =
//     V [rho_ki rho_lj - rho_li rho_kj + A01_ij A10_kl / 4] + d (kl <-> ij),
//     exact algebra; entries with a = b or c = d add exactly zero;
//   * <J^2> (ZJ2_0, PNAMP_ME2B.f90:1-152) in the ladder form
//     J.J = Jz Jz + (J+ J- + J- J+)/2, E_kin, Tr rho per species (ZN_0).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../include/pnamp_gpu.h"
#include "pnamp_dev.h"
#include "plain_gpu.h"

namespace {

using namespace pnamp_dev;

constexpr int NOBS = PNAMP_NOBS;
constexpr int NOG_E2 = NOBS + 10;  // per gauge point: observables, then tr_s(Q_mu rho), then orbit traces

template <int N>
struct PlainCfg {
  static constexpr int NW = (3 * N + 31) / 32;     // waves: 32 stacked rows of [T; A V1; A* U1] each
  static constexpr int NTH = 64 * NW;
  static constexpr int HC = N / 2;                 // columns per thread (half a row)
  static constexpr int NU = N * (2 * N - 1);       // strict upper entries of the 2N x 2N matrix
  static constexpr int SL = (NU + NTH - 1) / NTH;  // Pfaffian slots per thread
  // waves per SIMD the register budgets are fitted to: the Pfaffian kernel's
  // slots grow as N^2 / NTH, the gauge kernel's register columns as N / 2;
  // the gauge kernel of N = 64 (133 KB of LDS) runs one workgroup per CU
  static constexpr int PF_WAVES = N <= 40 ? 4 : (N <= 48 ? 3 : 2);
  static constexpr int GAUGE_WAVES = N <= 48 ? 3 : 2;
};

// ------------------------------------------------------------------ setup
struct PSetup {
  int n, h, nalp, ngam, npack, NE, u0;
  const double *alp, *gam, *rtab;
  const int *pk_a, *pk_b, *blk_off, *blk_sz, *blk_base, *twom;
  const int *q0, *q1;
  const double *U, *V, *Z;
  double2 *hoist;  // [unit - u0][7][n*n]: X_p, X_n, Y_p, Y_n, R~V1, R~*U1 (column-major), M0 (row-major)
};

constexpr int RPK = 256;  // packed rotation entries (per species) held in LDS

__global__ __launch_bounds__(256) void k_plain_setup(PSetup p) {
  extern __shared__ double2 lds[];
  __shared__ double2 sR[RPK];
  __shared__ short m_off[32], m_sz[32], m_base[32];
  const int tid = threadIdx.x, n = p.n, h = p.h, nn = n * n;
  const int ul = blockIdx.x, u = p.u0 + ul, ip = u / p.NE, iang = u - ip * p.NE;
  const int ig = iang % p.ngam, ia = (iang / p.ngam) % p.nalp, ib = iang / (p.ngam * p.nalp);
  const double alp = p.alp[ia], gam = p.gam[ig];
  const int q0 = p.q0[ip], q1 = p.q1[ip];
  const double *V0 = p.V + (size_t)q0 * nn, *U0 = p.U + (size_t)q0 * nn;
  const double *V1 = p.V + (size_t)q1 * nn, *U1 = p.U + (size_t)q1 * nn, *Z1 = p.Z + (size_t)q1 * nn;
  double2 *out = p.hoist + (size_t)ul * 7 * nn;
  double2 *sA = lds, *sB = lds + nn;  // R~V1, R~*U1; later R~ Z1 in sA
  for (int i = tid; i < h; i += 256) {
    m_off[i] = (short)p.blk_off[i];
    m_sz[i] = (short)p.blk_sz[i];
    m_base[i] = (short)p.blk_base[i];
  }
  // R(a, b) = exp(-i alp m_a) d(beta)_{ab} exp(-i gam m_b) (ROTATION_MAT, PNAMP_Proj.f90:83-136)
  for (int e = tid; e < p.npack; e += 256) {
    const int a = p.pk_a[e], b = p.pk_b[e];
    sR[e] = cscale(p.rtab[(size_t)ib * p.npack + e], cexpi(-alp * 0.5 * p.twom[a] - gam * 0.5 * p.twom[b]));
  }
  __syncthreads();
  // R~ V1, R~* U1 (block-sparse rows)
  for (int e = tid; e < nn; e += 256) {
    const int i = e % n, j = e / n;
    const int sp = i >= h, il = i - sp * h;
    const int off = m_off[il], sz = m_sz[il], base = m_base[il];
    double2 a = cz(), b = cz();
    for (int k = 0; k < sz; ++k) {
      const double2 r = sR[base + (il - off) + k * sz];
      const int row = sp * h + off + k;
      rfma(a, V1[row + j * n], r);
      rfma(b, U1[row + j * n], cconj(r));
    }
    sA[e] = a;
    sB[e] = b;
    out[4 * nn + e] = a;
    out[5 * nn + e] = b;
  }
  __syncthreads();
  // X_s(a, c) = sum_{i in s} V0(i, a) (R~V1)(i, c), Y_s likewise with U0, R~*U1
  for (int e = tid; e < nn; e += 256) {
    const int a = e % n, c = e / n;
    double2 x[2] = {cz(), cz()}, y[2] = {cz(), cz()};
    for (int s = 0; s < 2; ++s)
      for (int i = s * h; i < (s + 1) * h; ++i) {
        rfma(x[s], V0[i + a * n], sA[i + c * n]);
        rfma(y[s], U0[i + a * n], sB[i + c * n]);
      }
    out[e] = x[0];
    out[nn + e] = x[1];
    out[2 * nn + e] = y[0];
    out[3 * nn + e] = y[1];
  }
  __syncthreads();
  // M0 = (R~ Z1) R~^T
  for (int e = tid; e < nn; e += 256) {
    const int i = e % n, j = e / n;
    const int sp = i >= h, il = i - sp * h;
    const int off = m_off[il], sz = m_sz[il], base = m_base[il];
    double2 a = cz();
    for (int k = 0; k < sz; ++k) rfma(a, Z1[(sp * h + off + k) + j * n], sR[base + (il - off) + k * sz]);
    sA[e] = a;
  }
  __syncthreads();
  for (int e = tid; e < nn; e += 256) {
    const int j = e % n, i = e / n;  // row-major output: consecutive threads, consecutive j
    const int sp = j >= h, jl = j - sp * h;
    const int off = m_off[jl], sz = m_sz[jl], base = m_base[jl];
    double2 a = cz();
    for (int k = 0; k < sz; ++k) cfma(a, sA[i + (sp * h + off + k) * n], sR[base + (jl - off) + k * sz]);
    out[6 * nn + (size_t)i * n + j] = a;
  }
}

// ------------------------------------------------------------------ gauge point
struct PGauge {
  int n, h, nphi, NE, zval, nval, nl, lbytes, nob2, nog, u0, nwg;
  double cscale, vac_scale, snpfaf, e0;
  const int *q0, *q1, *pz, *pn;
  const double *U, *V, *Z, *det;
  const double2 *hoist;
  const double *mtab, *t1b;
  const uint4 *lofs;
  const double *lme;
  const double *q2;
  const int4 *orbp;
  double2 *gout;  // [unit - u0][nphi^2][nog]: observables per gauge point, not yet norm-weighted
  int *gst;       // [unit - u0][nphi^2]
  double2 *gpf;   // [unit - u0][nphi^2]: Pfaffian of the gauge point's skew matrix
};

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
  const unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffull), m), hi = __shfl_xor((unsigned)(v >> 32), m);
  return ((unsigned long long)hi << 32) | lo;
}

template <int N>
__host__ __device__ constexpr size_t gauge_misc_bytes() {
  // wave keys [2][NW] (8 B), reductions [NW][20] (8 B), pivot order [N]
  // (4 B): 864 B at N = 40, so three workgroups of n = 40 (52.7 KB) share a
  // CU's 160 KB
  return 16 * PlainCfg<N>::NW + 160 * PlainCfg<N>::NW + 4 * N;
}

// Pfaffian per gauge point, its own launch: little LDS (the four old rows,
// 16 n complex) and ~100 VGPRs, so many workgroups per CU hide the serial
// steps' barriers and reductions.
template <int N>
__global__ __launch_bounds__(PlainCfg<N>::NTH) __attribute__((amdgpu_waves_per_eu(PlainCfg<N>::PF_WAVES))) void k_plain_pf(PGauge p) {
  using C = PlainCfg<N>;
  constexpr int NTH = C::NTH, NW = C::NW, SL = C::SL;
  extern __shared__ double2 lds[];
  const int n = p.n, h = p.h, nn = n * n, N2 = 2 * n;
  unsigned long long *const wkey = (unsigned long long *)(lds + 8 * N2);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned L = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (L >= (unsigned)p.nwg) return;
  const int ng = p.nphi * p.nphi;
  const int g = (int)(L % (unsigned)ng), ul = (int)(L / (unsigned)ng), u = p.u0 + ul;
  const int ip = u / p.NE;
  const int q0 = p.q0[ip];
  const int lp = g / p.nphi, ln = g - lp * p.nphi;
  const double phip = (p.nphi == 1) ? 0.0 : (2.0 * M_PI / p.nphi) * lp;
  const double phin = (p.nphi == 1) ? 0.0 : (2.0 * M_PI / p.nphi) * ln;
  const double2 ep = cexpi(phip), en = cexpi(phin);
  const double2 *const H = p.hoist + (size_t)ul * 7 * nn;
  const double *const Z0 = p.Z + (size_t)q0 * nn;
  // ---------------------------------------------------------------- Pfaffian
  // S = cscale [[D M0 D, -1], [1, -Z0]] (Pfaf_Over, strict upper triangle)
  double2 pf = make_double2(1.0, 0.0);
  {
    const int NU = n * (N2 - 1);
    double2 sv[SL];
    int sij[SL];
    const double2 e2p = cmul(ep, ep), epn = cmul(ep, en), e2n = cmul(en, en);
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int e = s * NTH + tid;
      double2 v = cz();
      int ij = 0;
      if (e < NU) {
        // row from the bottom t: entries [t(t-1)/2, t(t+1)/2) are row N2-1-t
        int t = (int)((1.0 + sqrt(1.0 + 8.0 * (double)e)) * 0.5);
        while (t * (t - 1) / 2 > e) --t;
        while (t * (t + 1) / 2 <= e) ++t;
        const int i = N2 - 1 - t, j = i + 1 + (e - t * (t - 1) / 2);
        ij = i | (j << 8);
        if (j < n) {
          const int sidx = (i >= h) + (j >= h);
          v = cmul(sidx == 0 ? e2p : (sidx == 1 ? epn : e2n), H[6 * nn + (size_t)i * n + j]);
        }
        else if (i < n) v = make_double2((j - n == i) ? -1.0 : 0.0, 0.0);
        else v = make_double2(-Z0[(i - n) + (size_t)(j - n) * n], 0.0);
        v = cscale(p.cscale, v);
      }
      sv[s] = v;
      sij[s] = ij;
    }
    bool zero = false;
    double2 *const pbuf = lds;  // [2][4][N2]: old rows NR, NC, I1, I2 (R and AB are free here)
    // pivot key of an entry: |S|^2 with the low 14 mantissa bits replaced by
    // its row-major rank (largest magnitude, then the first in row-major order)
    auto pkey = [&](double2 v, int i, int j) -> unsigned long long {
      const double mag = v.x * v.x + v.y * v.y;
      return ((unsigned long long)__double_as_longlong(mag) & ~0x3FFFull) | (unsigned long long)(0x3FFF - (i * N2 + j));
    };
    // step 0's key and old rows 0, 1 (later steps get theirs from the previous update sweep)
    unsigned long long key = 0ull;
    {
      const int cnt = n * (N2 - 1);
      double2 *const rNR = pbuf, *const rNC = pbuf + N2;
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        if (s * NTH < cnt && s * NTH + tid < cnt) {
          const int i = sij[s] & 255, j = sij[s] >> 8;
          const double2 v = sv[s];
          const unsigned long long kk = pkey(v, i, j);
          key = kk > key ? kk : key;
          if (i == 0) rNR[j] = v;
          if (i == 1) rNC[j] = v;
          if (i == 0 && j == 1) rNC[0] = cneg(v);
        }
      }
    }
    for (int k = 0; k < n - 1; ++k) {
      const int NR = 2 * k, NC = NR + 1, m = N2 - NR;
      const int cnt = m * (m - 1) / 2, cnt2 = (m - 2) * (m - 3) / 2;
      double2 *const rNR = pbuf + (k & 1) * 4 * N2, *const rNC = rNR + N2, *const rI1 = rNR + 2 * N2,
                     *const rI2 = rNR + 3 * N2;
      double2 *const nNR = pbuf + ((k + 1) & 1) * 4 * N2, *const nNC = nNR + N2;  // old rows of step k + 1
      key = wave_umax64(key);
      if (lane == 0) wkey[(k & 1) * NW + wave] = key;
      __syncthreads();
      unsigned long long best = wkey[(k & 1) * NW];
      for (int w = 1; w < NW; ++w) {
        const unsigned long long o = wkey[(k & 1) * NW + w];
        best = o > best ? o : best;
      }
      const int lex = 0x3FFF - (int)(best & 0x3FFFull);
      const int I1 = lex / N2, I2 = lex - (lex / N2) * N2;
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        if (s * NTH < cnt && s * NTH + tid < cnt) {
          const int i = sij[s] & 255, j = sij[s] >> 8;
          const double2 v = sv[s];
          if (i == I1) rI1[j] = v;
          if (j == I1) rI1[i] = cneg(v);
          if (i == I2) rI2[j] = v;
          if (j == I2) rI2[i] = cneg(v);
        }
      }
      __syncthreads();
      const double2 ss = rI1[I2];  // S(I1, I2), the pivot after the exchanges
      if (!(ss.x * ss.x + ss.y * ss.y > 1.0e-26)) {  // |pivot| <= 1e-13
        zero = true;
        break;
      }
      const double phas = ((I1 == NR) ? -1.0 : 1.0) * ((I2 == NC) ? -1.0 : 1.0);
      pf = cscale(phas, cmul(pf, ss));
      const double2 inv = crecip(ss);
      // composed transposition pi = (NR I1)(NC I2) on the trailing indices:
      // pi(I1) = NR, pi(I2) = NC (NR if I1 = NC); S'(x, y) = S(pi x, pi y).
      // The updated entries are step k + 1's trailing matrix: their key and
      // its old rows NR + 2, NR + 3 are taken in the same sweep.
      const int pI2 = (I1 == NC) ? NR : NC;
      key = 0ull;
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        if (s * NTH < cnt2 && s * NTH + tid < cnt2) {
          const int x = sij[s] & 255, y = sij[s] >> 8;
          const int px = (x == I1) ? NR : ((x == I2) ? pI2 : x);
          const int py = (y == I1) ? NR : ((y == I2) ? pI2 : y);
          double2 base = sv[s];
          if (px != x) base = (px == NR ? rNR : rNC)[py];
          else if (py != y) base = cneg((py == NR ? rNR : rNC)[px]);
          // S(i,j) += (S'(NC,i) S'(NR,j) - S'(NR,i) S'(NC,j)) / S'(NR,NC)
          double2 t = cmul(rI2[px], rI1[py]);
          cfms(t, rI1[px], rI2[py]);
          cfma(base, t, inv);
          sv[s] = base;
          const unsigned long long kk = pkey(base, x, y);
          key = kk > key ? kk : key;
          if (x == NR + 2) nNR[y] = base;
          if (x == NR + 3) nNC[y] = base;
          if (x == NR + 2 && y == NR + 3) nNC[NR + 2] = cneg(base);
        }
      }
    }
    if (zero) pf = cz();
    else if (tid == 0) pf = cmul(pf, sv[0]);  // entry 0 is S(2n-2, 2n-1)
    if (tid == 0) p.gpf[L] = pf;
  }

}

template <int N, bool E2>
__global__ __launch_bounds__(PlainCfg<N>::NTH) __attribute__((amdgpu_waves_per_eu(PlainCfg<N>::GAUGE_WAVES))) void k_plain_gauge(PGauge p) {
  using C = PlainCfg<N>;
  constexpr int NTH = C::NTH, NW = C::NW, HC = C::HC;
  extern __shared__ double2 lds[];
  const int n = p.n, h = p.h, nn = n * n, ld = n + 1;
  double2 *const R = lds;              // rho, column-major with leading dimension n + 1
  double2 *const AB = lds + n * ld;    // A01 (i < j at i + j n), A10 / 4 (i < j at j + i n)
  char *const misc = (char *)(AB + nn);
  unsigned long long *const wkey = (unsigned long long *)misc;
  double *const red = (double *)(misc + 16 * NW);
  unsigned *const kinv = (unsigned *)(red + 20 * NW);  // [N]: solution index of each pivot column, times 8 n
  double2 *const rowb = AB;                        // Gauss-Jordan row buffers [2][N] (the A-block is free then)

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware order: the nphi^2 workgroups of one unit stay on one XCD (and
  // its L2, where the unit's setup matrices are); gridDim.x is a multiple of 8
  const unsigned L = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (L >= (unsigned)p.nwg) return;
  const int ng = p.nphi * p.nphi;
  const int g = (int)(L % (unsigned)ng), ul = (int)(L / (unsigned)ng), u = p.u0 + ul;
  const int ip = u / p.NE;
  const int q0 = p.q0[ip], q1 = p.q1[ip];
  const int lp = g / p.nphi, ln = g - lp * p.nphi;
  const double phip = (p.nphi == 1) ? 0.0 : (2.0 * M_PI / p.nphi) * lp;
  const double phin = (p.nphi == 1) ? 0.0 : (2.0 * M_PI / p.nphi) * ln;
  const double2 ep = cexpi(phip), en = cexpi(phin);
  const double2 *const H = p.hoist + (size_t)ul * 7 * nn;
  const double *const Z0 = p.Z + (size_t)q0 * nn;

  for (int a = tid; a < N; a += NTH) kinv[a] = 0u;  // padding columns (c >= n) read column 0

  // ---------------------------------------------------------------- Gauss-Jordan
  // stacked row r: T (r < n), A V1 (n <= r < 2n), A* U1 (2n <= r < 3n); lanes
  // l and l + 32 of a wave hold the two halves of one row
  const int r = 32 * wave + (lane & 31), hf = lane >> 5;
  const int kind = (r < n) ? 0 : (r < 2 * n) ? 1 : (r < 3 * n) ? 2 : 3;
  const int ri = r - (kind < 3 ? kind : 0) * n;
  double2 col[HC];
  {
    const double2 *Xp = H, *Xn = H + nn, *Yp = H + 2 * nn, *Yn = H + 3 * nn, *RV1 = H + 4 * nn, *RcU1 = H + 5 * nn;
    const double2 es = (ri >= h) ? en : ep;
    const double2 ph = (kind == 1) ? es : cconj(es);
#pragma unroll
    for (int cc = 0; cc < HC; ++cc) {
      const int c = hf * HC + cc;
      double2 v = cz();
      if (c < n && kind < 3) {
        const size_t o = (size_t)ri + (size_t)c * n;
        if (kind == 0) {
          v = cmul(ep, Xp[o]);
          cfma(v, en, Xn[o]);
          cfma(v, cconj(ep), Yp[o]);
          cfma(v, cconj(en), Yn[o]);
        } else {
          v = cmul(ph, (kind == 1 ? RV1 : RcU1)[o]);
        }
      }
      col[cc] = v;
    }
  }
  int info = 0;
  unsigned long long used = 0ull;
  for (int k = 0; k < n; ++k) {
    double2 *const b = rowb + (k & 1) * N;
    if (r == k) {
#pragma unroll
      for (int cc = 0; cc < HC; ++cc) b[hf * HC + cc] = col[cc];
    }
    __syncthreads();
    // pivot: largest |re| + |im| among unused columns, ties to the smaller
    // column (one DPP maximum; a NaN entry counts as zero and still leaves an
    // unused column as the pivot)
    unsigned pk = 0u;
    if (lane < n && !((used >> lane) & 1ull)) pk = pivot_key<6>(cabs1(b[lane]), lane);
    pk = wave_umax(pk);
    const int cs = 63 - (int)(pk & 63u);
    used |= 1ull << cs;
    const double2 piv = b[cs];
    if (piv.x == 0.0 && piv.y == 0.0 && !info) info = k + 1;
    const double2 ipv = crecip(piv);
    const int hs = cs / HC, cl = cs - hs * HC;  // owner half and register column (uniform)
    double2 x = cz();
#pragma unroll
    for (int cc = 0; cc < HC; ++cc)
      if (cc == cl) x = col[cc];
    double2 xlo, xhi;
    halves2(x, xlo, xhi);
    const double2 xp = cmul(hs ? xhi : xlo, ipv);  // pivot-column entry of this row (owner half), scaled
    if (tid == 0) kinv[cs] = (unsigned)(k * n) * 8u;
#pragma unroll
    for (int cc = 0; cc < HC; ++cc) {
      if (cc % 4 == 0) __builtin_amdgcn_sched_barrier(0);  // bounds the LDS reads in flight
      const double2 f = b[hf * HC + cc];
      if (cc == cl && hf == hs) col[cc] = xp;
      else cfms(col[cc], f, xp);
    }
  }
  // solution W(row, k) = col[cc] for the register column cc with pivot order k

  // ---------------------------------------------------------------- densities
  __syncthreads();  // the row buffers in the A-block are read no more
  for (int e = tid; e < nn; e += NTH) AB[e] = cz();
  __syncthreads();
  {
    const double *V0 = p.V + (size_t)q0 * nn, *U0 = p.U + (size_t)q0 * nn;
    const bool w1 = kind == 1, w2 = kind == 2;
    // outputs in pairs of columns j (n is even); U0 only in waves with W1 rows
    auto products = [&](auto withu) {
      constexpr bool WU = decltype(withu)::value;
#pragma unroll 1
      for (int j0 = 0; j0 < n; j0 += 2) {
        double2 a1[2] = {cz(), cz()}, a2[2] = {cz(), cz()};
#pragma unroll
        for (int cc = 0; cc < HC; ++cc) {
          const double2 w = col[cc];
          const unsigned o = kinv[hf * HC + cc] + (unsigned)j0 * 8u;
          const double2 v01 = *(const double2 *)((const char *)V0 + o);
          rfma(a1[0], v01.x, w);
          rfma(a1[1], v01.y, w);
          if constexpr (WU) {
            const double2 u01 = *(const double2 *)((const char *)U0 + o);
            rfma(a2[0], u01.x, w);
            rfma(a2[1], u01.y, w);
          }
          // at most four columns' loads in flight: the accumulators pass
          // through an empty asm that also orders memory, so the next
          // columns' loads wait for these multiply-adds (the scheduler would
          // otherwise issue all loads first and spill the W row)
          if (cc % 4 == 3) {
            __asm__ volatile("" : "+v"(a1[0].x), "+v"(a1[0].y), "+v"(a1[1].x), "+v"(a1[1].y) : : "memory");
            if constexpr (WU) __asm__ volatile("" : "+v"(a2[0].x), "+v"(a2[0].y), "+v"(a2[1].x), "+v"(a2[1].y) : : "memory");
          }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          double2 lo, hi;
          halves2(a1[t], lo, hi);
          a1[t] = cadd(lo, hi);
          if constexpr (WU) {
            halves2(a2[t], lo, hi);
            a2[t] = cadd(lo, hi);
          }
          const int j = j0 + t;
          if (hf == t) {
            if (w1) {
              R[ri + j * ld] = a1[t];  // rho
              if (ri != j) {           // kappa10 -> A10 / 4
                const double s = (ri < j) ? 0.25 : -0.25;
                double *dst = (double *)&AB[(ri < j ? j : ri) + (ri < j ? ri : j) * n];
                atomicAdd(dst, s * a2[t].x);
                atomicAdd(dst + 1, s * a2[t].y);
              }
            } else if (w2 && ri != j) {  // kappa01 = -a1 -> A01
              const double s = (ri < j) ? -1.0 : 1.0;
              double *dst = (double *)&AB[(ri < j ? ri : j) + (ri < j ? j : ri) * n];
              atomicAdd(dst, s * a1[t].x);
              atomicAdd(dst + 1, s * a1[t].y);
            }
          }
        }
      }
    };
    if (__any(w1)) products(std::true_type{});
    else if (__any(w2)) products(std::false_type{});
  }
  __syncthreads();

  // ---------------------------------------------------------------- observables
  // per element rho(a, b): E_kin t1(b, a); <J^2> rho part -m_a m_b rho(a,b) rho(b,a)
  // - c_a c_b rho(a,b) rho(b-1,a-1) + j_a(j_a+1) d_ab; kappa part
  // (A10/4)(a,b) [m_a m_b A01(a,b) + c_a c_{b+1}/2 A01(a-1,b+1) + c_{a+1} c_b/2 A01(a+1,b-1)]
  // (c_a = <m_a+1|J+|m_a>, 0 for the first state of an orbit, m = j .. -j);
  // traces t_z = sum m_a rho_aa, t_+ = sum c_a rho(a, a-1), t_- = sum c_{a+1} rho(a, a+1)
  double2 ek = cz(), aj = cz(), tz = cz(), tpl = cz(), tmi = cz(), zp = cz(), zn = cz();
  auto a01 = [&](int x, int y) -> double2 {
    return x < y ? AB[x + y * n] : (x > y ? cneg(AB[y + x * n]) : cz());
  };
  for (int e = tid; e < nn; e += NTH) {
    const int a = e % n, b = e / n;
    const double2 rab = R[a + b * ld], rba = R[b + a * ld];
    const int al = a >= h ? a - h : a, bl = b >= h ? b - h : b;
    const double ma = p.mtab[al], mb = p.mtab[bl], ca = p.mtab[h + al], cb = p.mtab[h + bl];
    rfma(ek, p.t1b[b + (size_t)a * n], rab);
    double2 q = cscale(ma * mb, rba);
    if (ca != 0.0 && cb != 0.0) rfma(q, ca * cb, R[(b - 1) + (a - 1) * ld]);
    cfms(aj, rab, q);
    if (a == b) {
      rfma(aj, p.mtab[2 * h + al], rab);
      rfma(tz, ma, rab);
      if (a < h) zp = cadd(zp, rab);
      else zn = cadd(zn, rab);
    }
    if (a == b + 1) rfma(tpl, ca, rab);
    if (b == a + 1) rfma(tmi, cb, rab);
    if (a != b) {
      const double2 l4 = (a < b) ? AB[b + a * n] : cneg(AB[a + b * n]);
      double2 s = cscale(ma * mb, a01(a, b));
      if (ca != 0.0 && b + 1 < n) {
        const double cb1 = p.mtab[h + (b + 1 >= h ? b + 1 - h : b + 1)];
        if (cb1 != 0.0) rfma(s, 0.5 * ca * cb1, a01(a - 1, b + 1));
      }
      if (cb != 0.0 && a + 1 < n) {
        const double ca1 = p.mtab[h + (a + 1 >= h ? a + 1 - h : a + 1)];
        if (ca1 != 0.0) rfma(s, 0.5 * ca1 * cb, a01(a + 1, b - 1));
      }
      cfma(aj, l4, s);
    }
  }
  // two-body energy, regrouped list
  double2 e2 = cz();
  {
    const char *lb = (const char *)lds;
    const int shift = p.lbytes ? 0 : 4;
    auto at = [&](unsigned o) { return *(const double2 *)(lb + (o << shift)); };
    for (int q = tid; q < p.nl; q += NTH) {
      const uint4 o = p.lofs[q];
      const double me = p.lme[q];
      double2 t = cmul(at(o.x & 0xffffu), at(o.x >> 16));
      cfms(t, at(o.y & 0xffffu), at(o.y >> 16));
      cfma(t, at(o.z & 0xffffu), at(o.z >> 16));
      rfma(e2, me, t);
    }
  }
  double v[16] = {e2.x, e2.y, ek.x, ek.y, aj.x, aj.y, tz.x, tz.y, tpl.x, tpl.y, tmi.x, tmi.y, zp.x, zp.y, zn.x, zn.y};
#pragma unroll
  for (int k = 0; k < 16; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 16; ++k) red[wave * 20 + k] = v[k];
  __syncthreads();
  double2 *const go = p.gout + (size_t)L * p.nog;
  if (tid == 0) {
    double t[16];
    for (int k = 0; k < 16; ++k) t[k] = 0.0;
    for (int w = 0; w < NW; ++w)
      for (int k = 0; k < 16; ++k) t[k] += red[w * 20 + k];
    const double2 E2v = make_double2(t[0], t[1]), EK = make_double2(t[2], t[3]), A = make_double2(t[4], t[5]);
    const double2 TZ = make_double2(t[6], t[7]), TP = make_double2(t[8], t[9]), TM = make_double2(t[10], t[11]);
    const double2 ZP = make_double2(t[12], t[13]), ZN = make_double2(t[14], t[15]);
    const double2 J2 = cadd(A, cadd(cmul(TZ, TZ), cmul(TP, TM)));
    go[0] = make_double2(1.0, 0.0);
    go[1] = ZP;
    go[2] = ZN;
    go[3] = EK;
    go[4] = cadd(cadd(E2v, EK), make_double2(p.e0, 0.0));
    go[5] = J2;
    p.gst[L] = info;
  }
  if (E2) {
    // tr_s(Q_mu rho) = sum_{i, j in s} Q_mu(i, j) rho(j, i) (ZQ2, PNAMP_ME1B.f90:52-75)
    __syncthreads();
    double q[20];
#pragma unroll
    for (int k = 0; k < 20; ++k) q[k] = 0.0;
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
      for (int rr = tid; rr < h * h; rr += NTH) {
        const int i = sp * h + rr % h, j = sp * h + rr / h;
        const double2 rji = R[j + i * ld];
#pragma unroll
        for (int mu = 0; mu < 5; ++mu) {
          const double qv = p.q2[(size_t)mu * nn + i + (size_t)j * n];
          q[sp * 10 + 2 * mu] = fma(qv, rji.x, q[sp * 10 + 2 * mu]);
          q[sp * 10 + 2 * mu + 1] = fma(qv, rji.y, q[sp * 10 + 2 * mu + 1]);
        }
      }
#pragma unroll
    for (int k = 0; k < 20; ++k)
      for (int off = 32; off > 0; off >>= 1) q[k] += __shfl_xor(q[k], off);
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 20; ++k) red[wave * 20 + k] = q[k];
    __syncthreads();
    if (tid < 10) {
      double re = 0.0, im = 0.0;
      for (int w = 0; w < NW; ++w) {
        re += red[w * 20 + 2 * tid];
        im += red[w * 20 + 2 * tid + 1];
      }
      go[NOBS + tid] = make_double2(re, im);
    }
    // orbit-pair traces sum_m rho(k_a(m), k_b(m)) (ZRho1B_Angles, PNAMP_Rho1B.f90:13-33)
    for (int o = tid; o < p.nob2; o += NTH) {
      const int4 op = p.orbp[o];
      double2 t = cz();
      for (int m = 0; m < op.w; ++m) t = cadd(t, R[(op.y + m) + (op.z + m) * ld]);
      go[NOG_E2 + o] = t;
    }
  }
}

// ------------------------------------------------------------------ gauge sum
struct PSum {
  int ng, nog, nob2, u0, nunits, NE, nphi, zval, nval;
  double vac_scale, snpfaf;
  const int *q0, *q1, *pz, *pn;
  const double *det;
  const double2 *gout, *gpf;
  const int *gst;
  double2 *zres, *ze2, *zrho;
  int *status;
};

// Fomenko sum over the gauge points of one unit: the norm weight
// znorm = <0|Phi0><0|Phi1> Pf e^{-i (phi_p Z + phi_n N)} times each observable
// (PNAMP_Kernel3D.f90:262-387), divided by nphi^2.
__global__ __launch_bounds__(64) void k_plain_sum(PSum p) {
  __shared__ double2 zn[256];
  const int ul = blockIdx.x, u = p.u0 + ul, ip = u / p.NE;
  const double inv = 1.0 / (double)p.ng;
  const double vac = p.snpfaf * (p.det[p.q0[ip]] * p.vac_scale) * (p.det[p.q1[ip]] * p.vac_scale);
  const double zv = p.pz ? p.pz[ip] : p.zval, nv = p.pn ? p.pn[ip] : p.nval;
  for (int g = threadIdx.x; g < p.ng; g += 64) {
    const int lp = g / p.nphi, ln = g - lp * p.nphi;
    const double phip = (p.nphi == 1) ? 0.0 : (2.0 * M_PI / p.nphi) * lp;
    const double phin = (p.nphi == 1) ? 0.0 : (2.0 * M_PI / p.nphi) * ln;
    zn[g] = cmul(cscale(vac, p.gpf[(size_t)ul * p.ng + g]), cexpi(-phip * zv - phin * nv));
  }
  __syncthreads();
  const double2 *gi = p.gout + (size_t)ul * p.ng * p.nog;
  for (int o = threadIdx.x; o < p.nog; o += 64) {
    double2 s = cz();
    for (int g = 0; g < p.ng; ++g) {
      if (o == 0) s = cadd(s, zn[g]);
      else cfma(s, zn[g], gi[(size_t)g * p.nog + o]);
    }
    s = cscale(inv, s);
    if (o < NOBS) p.zres[(size_t)u * NOBS + o] = s;
    else if (o < NOG_E2) p.ze2[(size_t)u * 10 + (o - NOBS)] = s;
    else p.zrho[(size_t)u * p.nob2 + (o - NOG_E2)] = s;
  }
  if (threadIdx.x == 0) {
    int st = 0;
    for (int g = 0; g < p.ng && !st; ++g) st = p.gst[(size_t)ul * p.ng + g];
    p.status[u] = st;
  }
}

// ------------------------------------------------------------------ host
// Size classes of the kernels (n <= N): sd 24, pf 40, f5pg9 44 -> 48,
// sdpf 64.  The gauge kernel's LDS is 16 (2 n^2 + n) bytes plus ~1 KB:
// 52.7 KB at n = 40 (three workgroups per CU), 62.6 KB at n = 44 (two),
// 133 KB at n = 64 (one).
constexpr int NCLASS = 6;
constexpr int CLASSES[NCLASS] = {16, 24, 32, 40, 48, 64};

int class_of(int n) {
  for (int c : CLASSES)
    if (n <= c) return c;
  return 0;
}

template <int N>
size_t gauge_lds(int n) {
  return sizeof(double2) * ((size_t)n * (n + 1) + (size_t)n * n) + gauge_misc_bytes<N>();
}

template <int N>
size_t pf_lds(int n) {
  return sizeof(double2) * 16 * (size_t)n + 16 * PlainCfg<N>::NW;
}

template <int N, bool E2>
hipError_t prepare_nc(int n) {
  return hipFuncSetAttribute((const void *)k_plain_gauge<N, E2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)gauge_lds<N>(n));
}

template <int N>
hipError_t launch_gauge(bool e2, unsigned grid, hipStream_t s, const PGauge &gp) {
  hipLaunchKernelGGL((k_plain_pf<N>), dim3(grid), dim3(PlainCfg<N>::NTH), pf_lds<N>(gp.n), s, gp);
  const size_t lds = gauge_lds<N>(gp.n);
  if (e2) hipLaunchKernelGGL((k_plain_gauge<N, true>), dim3(grid), dim3(PlainCfg<N>::NTH), lds, s, gp);
  else hipLaunchKernelGGL((k_plain_gauge<N, false>), dim3(grid), dim3(PlainCfg<N>::NTH), lds, s, gp);
  return hipGetLastError();
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

bool plain_supported(int n) { return n >= 8 && !(n & 1) && class_of(n) != 0; }

size_t plain_unit_bytes(const PlainArgs &a) {
  const size_t nn = (size_t)a.n * a.n, ng = (size_t)a.nphi * a.nphi;
  const size_t nog = a.q2 ? (size_t)NOG_E2 + a.nob2 : (size_t)NOBS;
  return align256(sizeof(double2) * 7 * nn) + align256(sizeof(double2) * ng * nog) + align256(sizeof(int) * ng) +
         align256(sizeof(double2) * ng);
}

hipError_t plain_prepare(int n, bool e2) {
  const hipError_t e = hipFuncSetAttribute((const void *)k_plain_setup, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(sizeof(double2) * 2 * (size_t)n * n));
  if (e != hipSuccess) return e;
  switch (class_of(n)) {
    case 16: return e2 ? prepare_nc<16, true>(n) : prepare_nc<16, false>(n);
    case 24: return e2 ? prepare_nc<24, true>(n) : prepare_nc<24, false>(n);
    case 32: return e2 ? prepare_nc<32, true>(n) : prepare_nc<32, false>(n);
    case 40: return e2 ? prepare_nc<40, true>(n) : prepare_nc<40, false>(n);
    case 48: return e2 ? prepare_nc<48, true>(n) : prepare_nc<48, false>(n);
    case 64: return e2 ? prepare_nc<64, true>(n) : prepare_nc<64, false>(n);
    default: return hipErrorInvalidValue;
  }
}

hipError_t plain_run(const PlainArgs &a, int u0, int nu, void *work, hipStream_t s) {
  const int N = class_of(a.n);
  if (!N || nu <= 0) return hipErrorInvalidValue;
  const size_t nn = (size_t)a.n * a.n, ng = (size_t)a.nphi * a.nphi;
  const bool e2 = a.q2 != nullptr;
  const int nog = e2 ? NOG_E2 + a.nob2 : NOBS;
  // work buffer: all units' setup matrices, then all gauge outputs, then all gauge status
  char *w = (char *)work;
  double2 *hoist = (double2 *)w;
  double2 *gout = (double2 *)(w + align256(sizeof(double2) * 7 * nn * nu));
  int *gst = (int *)((char *)gout + align256(sizeof(double2) * ng * nog * nu));
  double2 *gpf = (double2 *)((char *)gst + align256(sizeof(int) * ng * nu));
  if (ng > 256) return hipErrorInvalidValue;  // nphi <= 16

  PSetup sp;
  sp.n = a.n; sp.h = a.h; sp.nalp = a.nalp; sp.ngam = a.ngam; sp.npack = a.npack; sp.NE = a.NE; sp.u0 = u0;
  sp.alp = a.alp; sp.gam = a.gam; sp.rtab = a.rtab; sp.pk_a = a.pk_a; sp.pk_b = a.pk_b;
  sp.blk_off = a.blk_off; sp.blk_sz = a.blk_sz; sp.blk_base = a.blk_base; sp.twom = a.twom;
  sp.q0 = a.q0; sp.q1 = a.q1; sp.U = a.U; sp.V = a.V; sp.Z = a.Z; sp.hoist = hoist;
  hipLaunchKernelGGL(k_plain_setup, dim3(nu), dim3(256), sizeof(double2) * 2 * nn, s, sp);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;

  PGauge gp;
  gp.n = a.n; gp.h = a.h; gp.nphi = a.nphi; gp.NE = a.NE; gp.zval = a.zval; gp.nval = a.nval;
  gp.nl = a.nl; gp.lbytes = a.lbytes; gp.nob2 = a.nob2; gp.nog = nog; gp.u0 = u0;
  gp.nwg = (int)(ng * nu);
  gp.cscale = a.cscale; gp.vac_scale = a.vac_scale; gp.snpfaf = a.snpfaf; gp.e0 = a.e0;
  gp.q0 = a.q0; gp.q1 = a.q1; gp.pz = a.pz; gp.pn = a.pn;
  gp.U = a.U; gp.V = a.V; gp.Z = a.Z; gp.det = a.det; gp.hoist = hoist;
  gp.mtab = a.mtab; gp.t1b = a.t1b; gp.lofs = a.lofs; gp.lme = a.lme; gp.q2 = a.q2; gp.orbp = a.orbp;
  gp.gout = gout; gp.gst = gst; gp.gpf = gpf;
  const unsigned grid = (unsigned)((ng * nu + 7) / 8 * 8);
  switch (N) {
    case 16: err = launch_gauge<16>(e2, grid, s, gp); break;
    case 24: err = launch_gauge<24>(e2, grid, s, gp); break;
    case 32: err = launch_gauge<32>(e2, grid, s, gp); break;
    case 40: err = launch_gauge<40>(e2, grid, s, gp); break;
    case 48: err = launch_gauge<48>(e2, grid, s, gp); break;
    default: err = launch_gauge<64>(e2, grid, s, gp); break;
  }
  if (err != hipSuccess) return err;

  PSum sm;
  sm.ng = (int)ng; sm.nog = nog; sm.nob2 = a.nob2; sm.u0 = u0; sm.nunits = nu; sm.NE = a.NE; sm.nphi = a.nphi;
  sm.zval = a.zval; sm.nval = a.nval; sm.vac_scale = a.vac_scale; sm.snpfaf = a.snpfaf;
  sm.q0 = a.q0; sm.q1 = a.q1; sm.pz = a.pz; sm.pn = a.pn; sm.det = a.det;
  sm.gout = gout; sm.gpf = gpf; sm.gst = gst; sm.zres = a.zres; sm.ze2 = a.ze2; sm.zrho = a.zrho; sm.status = a.status;
  hipLaunchKernelGGL(k_plain_sum, dim3(nu), dim3(64), 0, s, sm);
  return hipGetLastError();
}
