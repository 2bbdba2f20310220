/*
 * pnamp_oracle.c -- CPU FP64 restatement of the reference PNAMP projected-kernel
 * path (jmyao17/SCCM).  TEST INFRASTRUCTURE ONLY: this library is the checker
 * for the HIP product path and the CPU baseline timed beside it.  Nothing in
 * the product (sccm_amd/, include/) links or calls it.
 *
 * Every function follows the reference routine named in its comment, in the
 * same call sequence (the LAPACK/BLAS calls of the reference are restated as
 * plain loops: ZGEMM, ZGETRF with |re|+|im| pivoting, ZGETRI, a determinant
 * for clingd, and a cyclic Jacobi eigensolver for ZGEEV on a real symmetric
 * argument and for sdiag).
 *
 * Pinning of each function (the checks live in tests/):
 *   wigner_d        closed form vs the textbook d^{1/2}, d^1 and unitarity/composition  (test_oracle_math)
 *   rotation_mat    R unitary, R(O1)R(O2)=R(O1 O2)                                    (test_oracle_math)
 *   zpfaffianf      Pf^2 = det, block-diagonal closed form, brute-force Pfaffian       (test_oracle_math)
 *   zbareover       closed form det(1+Z^T Z)^(-1/4); exact Fock-space vacuum overlap    (test_fock)
 *   pfaf_over/norm  exact Fock-space <Phi0|R e^{i phi N}|Phi1> (magnitude and sign)     (test_fock)
 *   onishi          |N| equals the Pfaffian norm; Ge76 golden kernels                   (test_golden_ge76)
 *   rokappa         exact Fock-space transition densities                              (test_fock)
 *   energy          sparse list (Method II) == field form (Method I); Fock-space <H>;  (test_fock)
 *                   Ge76 golden H^J (8 digits)                                         (test_golden_ge76)
 *   projection      plain: exact Fock-space P^J_{KK'} kernels; literal fold: Ge76
 *                   axial goldens (PNAMP_Kernel1D path, K=0)                          (test_golden_ge76)
 *   hw_solve        numpy eigh textbook HNDIAG                                         (test_oracle_math)
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;
#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* ---------------------------------------------------------------- BLAS-like */

/* C = alpha * op(A) * op(B) + beta * C, column-major, ta/tb in {'n','t'}
 * (ZGEMM with the loop order of the reference BLAS: column j of C is
 * accumulated from columns of A for op(A) = A, by dot products for A^T). */
static void zgemm(char ta, char tb, int m, int n, int k, zc alpha, const zc *A, int lda,
                  const zc *B, int ldb, zc beta, zc *C, int ldc) {
  for (int j = 0; j < n; ++j) {
    zc *c = C + (size_t)j * ldc;
    if (beta == 0)
      for (int i = 0; i < m; ++i) c[i] = 0;
    else if (beta != 1)
      for (int i = 0; i < m; ++i) c[i] *= beta;
    if (ta == 'n') {
      for (int l = 0; l < k; ++l) {
        zc b = (tb == 'n') ? B[IX(l, j, ldb)] : B[IX(j, l, ldb)];
        if (b == 0) continue;
        b *= alpha;
        const zc *a = A + (size_t)l * lda;
        for (int i = 0; i < m; ++i) c[i] += a[i] * b;
      }
    } else {
      for (int i = 0; i < m; ++i) {
        const zc *a = A + (size_t)i * lda;
        zc s = 0;
        if (tb == 'n') {
          const zc *bb = B + (size_t)j * ldb;
          for (int l = 0; l < k; ++l) s += a[l] * bb[l];
        } else {
          for (int l = 0; l < k; ++l) s += a[l] * B[IX(j, l, ldb)];
        }
        c[i] += alpha * s;
      }
    }
  }
}

/* ZGETRF: LU with partial pivoting (pivot = max |re|+|im|, first maximum). */
static int zgetrf(int n, zc *a, int lda, int *ipiv) {
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double m = cabs1(a[IX(j, j, lda)]);
    for (int i = j + 1; i < n; ++i) {
      double v = cabs1(a[IX(i, j, lda)]);
      if (v > m) { m = v; p = i; }
    }
    ipiv[j] = p;
    if (a[IX(p, j, lda)] != 0) {
      if (p != j)
        for (int k = 0; k < n; ++k) {
          zc t = a[IX(j, k, lda)]; a[IX(j, k, lda)] = a[IX(p, k, lda)]; a[IX(p, k, lda)] = t;
        }
      zc r = 1.0 / a[IX(j, j, lda)];
      for (int i = j + 1; i < n; ++i) a[IX(i, j, lda)] *= r;
    } else if (!info) {
      info = j + 1;
    }
    for (int k = j + 1; k < n; ++k) {
      zc t = a[IX(j, k, lda)];
      if (t != 0)
        for (int i = j + 1; i < n; ++i) a[IX(i, k, lda)] -= a[IX(i, j, lda)] * t;
    }
  }
  return info;
}

/* ZGETRI: inverse from the LU factors (solve A X = I column by column). */
static int zgetri(int n, zc *a, int lda, const int *ipiv) {
  zc *x = (zc *)calloc((size_t)n * n, sizeof(zc));
  for (int j = 0; j < n; ++j) {
    zc *col = x + (size_t)j * n;
    col[j] = 1.0;
    for (int i = 0; i < n; ++i) { /* apply row interchanges */
      int p = ipiv[i];
      if (p != i) { zc t = col[i]; col[i] = col[p]; col[p] = t; }
    }
    for (int i = 0; i < n; ++i) /* L y = P b (unit lower) */
      for (int k = 0; k < i; ++k) col[i] -= a[IX(i, k, lda)] * col[k];
    for (int i = n - 1; i >= 0; --i) { /* U x = y */
      for (int k = i + 1; k < n; ++k) col[i] -= a[IX(i, k, lda)] * col[k];
      col[i] /= a[IX(i, i, lda)];
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[IX(i, j, lda)] = x[IX(i, j, n)];
  free(x);
  return 0;
}

static int zinverse(int n, zc *a) {
  int *ipiv = (int *)malloc(sizeof(int) * n);
  int info = zgetrf(n, a, n, ipiv);
  if (!info) zgetri(n, a, n, ipiv);
  free(ipiv);
  return info;
}

/* clingd: complex determinant by LU. */
static zc zdet(int n, const zc *a_in) {
  zc *a = (zc *)malloc(sizeof(zc) * n * n);
  memcpy(a, a_in, sizeof(zc) * n * n);
  int *ipiv = (int *)malloc(sizeof(int) * n);
  zgetrf(n, a, n, ipiv);
  zc d = 1.0;
  for (int i = 0; i < n; ++i) {
    d *= a[IX(i, i, n)];
    if (ipiv[i] != i) d = -d;
  }
  free(a); free(ipiv);
  return d;
}

/* Cyclic Jacobi for a real symmetric matrix: eigenvalues ascending,
 * eigenvectors in the columns of v (may be NULL). */
static void jacobi_eig(int n, double *a, double *w, double *v) {
  double *vv = v ? v : (double *)malloc(sizeof(double) * n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) vv[IX(i, j, n)] = (i == j);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double x = a[IX(i, j, n)] * a[IX(i, j, n)];
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-30 * tot || off == 0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[IX(p, q, n)];
        if (fabs(apq) < 1e-300) continue;
        double app = a[IX(p, p, n)], aqq = a[IX(q, q, n)];
        double theta = (aqq - app) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        double c = 1 / sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = a[IX(k, p, n)], akq = a[IX(k, q, n)];
          a[IX(k, p, n)] = c * akp - s * akq;
          a[IX(k, q, n)] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = a[IX(p, k, n)], aqk = a[IX(q, k, n)];
          a[IX(p, k, n)] = c * apk - s * aqk;
          a[IX(q, k, n)] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = vv[IX(k, p, n)], vkq = vv[IX(k, q, n)];
          vv[IX(k, p, n)] = c * vkp - s * vkq;
          vv[IX(k, q, n)] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = a[IX(i, i, n)];
  /* sort ascending */
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (w[j] < w[m]) m = j;
    if (m != i) {
      double t = w[i]; w[i] = w[m]; w[m] = t;
      for (int k = 0; k < n; ++k) {
        double x = vv[IX(k, i, n)]; vv[IX(k, i, n)] = vv[IX(k, m, n)]; vv[IX(k, m, n)] = x;
      }
    }
  }
  if (!v) free(vv);
}

/* ---------------------------------------------------------------- Wigner d */

static double fact_tab[64];
static void init_fact(void) {
  if (fact_tab[0] == 1.0) return;
  fact_tab[0] = 1.0;
  for (int i = 1; i < 64; ++i) fact_tab[i] = fact_tab[i - 1] * i;
}

/* d^j_{m1 m2}(beta) = <j m1| exp(-i beta J_y) |j m2>, arguments doubled
 * (Wigner / Varshalovich convention; the reference's Wigner_smalld,
 * Codes/GCM/VAPHFB_ThreeJ.f90:590-652, with its m-arguments swapped).
 * Stands in for the missing dwignerI_gen ('h' and 'i'), see
 * Codes/PNAMP/PNAMP_Kernel3D.f90:6-14 for the convention note. */
double or_wigner_d(int tj, int tm1, int tm2, double beta) {
  init_fact();
  if (abs(tm1) > tj || abs(tm2) > tj || ((tj + tm1) & 1) || ((tj + tm2) & 1)) return 0.0;
  int jpm1 = (tj + tm1) / 2, jmm1 = (tj - tm1) / 2, jpm2 = (tj + tm2) / 2, jmm2 = (tj - tm2) / 2;
  double pre = sqrt(fact_tab[jpm1] * fact_tab[jmm1] * fact_tab[jpm2] * fact_tab[jmm2]);
  double c = cos(0.5 * beta), s = sin(0.5 * beta);
  int dm = (tm1 - tm2) / 2; /* m1 - m2 */
  int smin = dm < 0 ? -dm : 0;
  int smax = jpm2 < jmm1 ? jpm2 : jmm1;
  double sum = 0;
  for (int k = smin; k <= smax; ++k) {
    double den = fact_tab[jpm2 - k] * fact_tab[k] * fact_tab[dm + k] * fact_tab[jmm1 - k];
    double term = pow(c, tj - dm - 2 * k) * pow(s, dm + 2 * k) / den;
    sum += ((dm + k) & 1) ? -term : term;
  }
  return pre * sum;
}

/* Wigner 3j symbol (j1 j2 j3; m1 m2 m3), arguments doubled, Racah formula
 * (the reference's wignei / Wigner_3j, Codes/GCM/VAPHFB_ThreeJ.f90). */
double or_threej(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3) {
  init_fact();
  if (tm1 + tm2 + tm3 != 0) return 0.0;
  if (abs(tm1) > tj1 || abs(tm2) > tj2 || abs(tm3) > tj3) return 0.0;
  if (((tj1 + tm1) & 1) || ((tj2 + tm2) & 1) || ((tj3 + tm3) & 1)) return 0.0;
  if (tj3 > tj1 + tj2 || tj3 < abs(tj1 - tj2) || ((tj1 + tj2 + tj3) & 1)) return 0.0;
  int a = (tj1 + tj2 - tj3) / 2, b = (tj1 - tj2 + tj3) / 2, cc = (-tj1 + tj2 + tj3) / 2;
  double tri = fact_tab[a] * fact_tab[b] * fact_tab[cc] / fact_tab[(tj1 + tj2 + tj3) / 2 + 1];
  double pre = sqrt(tri * fact_tab[(tj1 + tm1) / 2] * fact_tab[(tj1 - tm1) / 2] * fact_tab[(tj2 + tm2) / 2] *
                    fact_tab[(tj2 - tm2) / 2] * fact_tab[(tj3 + tm3) / 2] * fact_tab[(tj3 - tm3) / 2]);
  int kmin = 0, kmax = a;
  int t1 = (tj3 - tj2 + tm1) / 2, t2 = (tj3 - tj1 - tm2) / 2;
  if (-t1 > kmin) kmin = -t1;
  if (-t2 > kmin) kmin = -t2;
  if ((tj1 - tm1) / 2 < kmax) kmax = (tj1 - tm1) / 2;
  if ((tj2 + tm2) / 2 < kmax) kmax = (tj2 + tm2) / 2;
  double sum = 0.0;
  for (int k = kmin; k <= kmax; ++k) {
    double den = fact_tab[k] * fact_tab[a - k] * fact_tab[(tj1 - tm1) / 2 - k] * fact_tab[(tj2 + tm2) / 2 - k] *
                 fact_tab[t1 + k] * fact_tab[t2 + k];
    sum += (k & 1) ? -1.0 / den : 1.0 / den;
  }
  int ph = (tj1 - tj2 - tm3) / 2;
  return ((ph & 1) ? -1.0 : 1.0) * pre * sum;
}

/* ---------------------------------------------------------------- reference routines */

typedef struct {
  int n;
  const int *t, *nq, *l, *twoj, *twom;
  int64_t nme;
  const int *ka, *kb, *kc, *kd; /* 1-based */
  const double *me;
  const double *t1b; /* n x n column-major */
  double e0;
  int nalp, nbet, ngam;
  const double *alp, *walp, *bet, *wbet, *gam, *wgam;
  int nphi, jmax, fold, norm_method;
  double cscale;
  int iscale;
  int zval, nval;
  int nthreads;
  /* two-body energy per mesh point: 0 the list sum HFB_Energy_COMPLEX
   * (Method II, what Kernel4Triaxial calls, PNAMP_Kernel3D.f90:316), 1 the
   * field form HFB_FIELD_COMPLEX + HFB_ENER_comp (Method I, the commented
   * alternative at PNAMP_Kernel3D.f90:309-310) */
  int energy_method;
  /* E2 (iE2 = 1): the one-body Q_{2mu} matrices, [5][n*n] real column-major
   * for mu = -2..2 (cME1B%Q2_2t .. Q22t of SingleParticleME,
   * Codes/PNAMP/PNAMP_ME1B.f90:138-216), or NULL */
  const double *q2;
} or_cfg;

/* ROTATION_MAT (Codes/PNAMP/PNAMP_Proj.f90:83-136). */
void or_rotation_mat(const or_cfg *c, double alp, double bet, double gam, zc *R) {
  int n = c->n;
  for (size_t i = 0; i < (size_t)n * n; ++i) R[i] = 0;
  for (int ia = 0; ia < n; ++ia) {
    zc za = cexp(-I * alp * (c->twom[ia] / 2.0));
    for (int ib = 0; ib < n; ++ib) {
      if (c->nq[ib] == c->nq[ia] && c->l[ib] == c->l[ia] && c->t[ib] == c->t[ia] &&
          c->twoj[ib] == c->twoj[ia]) {
        zc zg = cexp(-I * gam * (c->twom[ib] / 2.0));
        R[IX(ia, ib, n)] = za * zg * or_wigner_d(c->twoj[ia], c->twom[ia], c->twom[ib], bet);
      }
    }
  }
}

/* zbareover (PNAMP_Proj.f90:6-79) for real U, V: Z = V U^{-1}, Zc = conj(Z),
 * Det = exp(-sum log(1+w)/4), w = eigenvalues of Zc^T Zc (real symmetric, PSD). */
int or_zbareover(int n, const double *U, const double *V, zc *Zc, double *det) {
  zc *Ui = (zc *)malloc(sizeof(zc) * n * n);
  zc *Vz = (zc *)malloc(sizeof(zc) * n * n);
  for (size_t i = 0; i < (size_t)n * n; ++i) { Ui[i] = U[i]; Vz[i] = V[i]; }
  int info = zinverse(n, Ui);
  if (info) { free(Ui); free(Vz); return info; }
  zc *Z = (zc *)malloc(sizeof(zc) * n * n);
  zgemm('n', 'n', n, n, n, 1.0, Vz, n, Ui, n, 0.0, Z, n);
  for (size_t i = 0; i < (size_t)n * n; ++i) Zc[i] = conj(Z[i]);
  zgemm('t', 'n', n, n, n, 1.0, Zc, n, Zc, n, 0.0, Z, n);
  double *a = (double *)malloc(sizeof(double) * n * n), *w = (double *)malloc(sizeof(double) * n);
  for (size_t i = 0; i < (size_t)n * n; ++i) a[i] = creal(Z[i]);
  jacobi_eig(n, a, w, NULL);
  double s = 0;
  for (int i = 0; i < n; ++i) s += log(1.0 + w[i]);
  *det = exp(-s / 4.0);
  free(a); free(w); free(Ui); free(Vz); free(Z);
  return 0;
}

/* UVphiomega (PNAMP_Proj.f90:396-455): A = diag(e^{i phi_p} R_pp, e^{i phi_n} R_nn),
 * Ubar = A U1, Vbar = conj(A) V1. */
void or_uvphiomega(int n, double phip, double phin, const zc *R, const double *U1,
                   const double *V1, zc *Ubar, zc *Vbar, zc *A) {
  int h = n / 2;
  zc *Ac = (zc *)calloc((size_t)n * n, sizeof(zc));
  for (size_t i = 0; i < (size_t)n * n; ++i) A[i] = 0;
  zc pp = cexp(I * phip), pn = cexp(I * phin);
  for (int i = 0; i < h; ++i)
    for (int j = 0; j < h; ++j) {
      A[IX(i, j, n)] = pp * R[IX(i, j, n)];
      A[IX(i + h, j + h, n)] = pn * R[IX(i + h, j + h, n)];
      Ac[IX(i, j, n)] = conj(A[IX(i, j, n)]);
      Ac[IX(i + h, j + h, n)] = conj(A[IX(i + h, j + h, n)]);
    }
  zc *Uz = (zc *)malloc(sizeof(zc) * n * n), *Vz = (zc *)malloc(sizeof(zc) * n * n);
  for (size_t i = 0; i < (size_t)n * n; ++i) { Uz[i] = U1[i]; Vz[i] = V1[i]; }
  zgemm('n', 'n', n, n, n, 1.0, A, n, Uz, n, 0.0, Ubar, n);
  zgemm('n', 'n', n, n, n, 1.0, Ac, n, Vz, n, 0.0, Vbar, n);
  free(Ac); free(Uz); free(Vz);
}

/* ZExch (PNAMP_Proj.f90:356-392): exchange rows/columns i1 < i2 of an
 * upper-triangle-stored skew matrix.  0-based indices. */
static void zexch(zc *S, int lds, int N, int i1, int i2) {
  S[IX(i1, i2, lds)] = -S[IX(i1, i2, lds)];
  for (int i = 0; i < i1; ++i) {
    zc t = S[IX(i, i1, lds)]; S[IX(i, i1, lds)] = S[IX(i, i2, lds)]; S[IX(i, i2, lds)] = t;
  }
  for (int i = i2 + 1; i < N; ++i) {
    zc t = S[IX(i1, i, lds)]; S[IX(i1, i, lds)] = S[IX(i2, i, lds)]; S[IX(i2, i, lds)] = t;
  }
  for (int i = i1 + 1; i < i2; ++i) {
    zc t = S[IX(i1, i, lds)];
    S[IX(i1, i, lds)] = -S[IX(i, i2, lds)];
    S[IX(i, i2, lds)] = -t;
  }
}

/* ZPfaffianF (PNAMP_Proj.f90:218-350): block Parlett-Reid with a full search
 * of the remaining upper triangle for the pivot; 0 if a pivot is <= 1e-13. */
zc or_zpfaffianf(zc *S, int lds, int N) {
  if (N % 2) return 0;
  int NB = N / 2;
  if (NB == 1) return S[IX(0, 1, lds)];
  zc pf = 1.0;
  const double epsln = 1.0e-13;
  for (int IB = 1; IB <= NB - 1; ++IB) {
    int NR = 2 * IB - 2, NC = NR + 1; /* 0-based */
    double big = 0;
    int I1 = NR, I2 = NC;
    for (int i = NR; i <= N - 2; ++i)
      for (int j = i + 1; j <= N - 1; ++j) {
        double v = cabs(S[IX(i, j, lds)]);
        if (v > big) { big = v; I1 = i; I2 = j; }
      }
    double phas = 1.0;
    if (I1 == NR) phas = -phas;
    if (I2 == NC) phas = -phas;
    if (I1 != NR) zexch(S, lds, N, NR, I1);
    if (I2 != NC) zexch(S, lds, N, NC, I2);
    zc ss = S[IX(NR, NC, lds)];
    if (cabs(ss) > epsln) {
      for (int i = 2 * IB; i <= N - 2; ++i)
        for (int j = i + 1; j <= N - 1; ++j)
          S[IX(i, j, lds)] += (S[IX(NC, i, lds)] * S[IX(NR, j, lds)] - S[IX(NR, i, lds)] * S[IX(NC, j, lds)]) / ss;
      for (int i = 2 * IB; i <= N - 1; ++i) {
        S[IX(i, NR, lds)] = -S[IX(NC, i, lds)] / ss;
        S[IX(i, NC, lds)] = S[IX(NR, i, lds)] / ss;
      }
      if (IB >= 2)
        for (int j = 0; j < 2 * IB; ++j) {
          zc sw = S[IX(NR, j, lds)]; S[IX(NR, j, lds)] = S[IX(I1, j, lds)]; S[IX(I1, j, lds)] = sw;
          sw = S[IX(NC, j, lds)]; S[IX(NC, j, lds)] = S[IX(I2, j, lds)]; S[IX(I2, j, lds)] = sw;
        }
    } else {
      return 0;
    }
    pf = pf * ss * phas;
  }
  return pf * S[IX(N - 2, N - 1, lds)];
}

/* Pfaf_Over (PNAMP_Proj.f90:142-211). */
int or_pfaf_over(int n, const zc *ZM0, const zc *Ubar, const zc *Vbar, double cscale, zc *pf) {
  zc *Ui = (zc *)malloc(sizeof(zc) * n * n), *Vc = (zc *)malloc(sizeof(zc) * n * n);
  for (size_t i = 0; i < (size_t)n * n; ++i) { Ui[i] = conj(Ubar[i]); Vc[i] = conj(Vbar[i]); }
  int info = zinverse(n, Ui);
  if (info) { free(Ui); free(Vc); return info; }
  zc *M1 = (zc *)malloc(sizeof(zc) * n * n);
  zgemm('n', 'n', n, n, n, 1.0, Vc, n, Ui, n, 0.0, M1, n);
  int N2 = 2 * n;
  zc *S = (zc *)calloc((size_t)N2 * N2, sizeof(zc));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      S[IX(j, i, N2)] = M1[IX(j, i, n)];
      S[IX(j + n, i + n, N2)] = -ZM0[IX(j, i, n)];
    }
  for (int i = 0; i < n; ++i) { S[IX(i, i + n, N2)] = -1.0; S[IX(i + n, i, N2)] = 1.0; }
  for (size_t i = 0; i < (size_t)N2 * N2; ++i) S[i] *= cscale;
  *pf = or_zpfaffianf(S, N2, N2);
  free(S); free(M1); free(Ui); free(Vc);
  return 0;
}

/* Onishi (Codes/PNAMP/PNAMP_Onishi.f90:1-48): sqrt(det(Ubar^T U0 + Vbar^T V0)),
 * principal branch like cdsqrt. */
zc or_onishi(int n, const double *U0, const double *V0, const zc *Ubar, const zc *Vbar) {
  zc *U0z = (zc *)malloc(sizeof(zc) * n * n), *V0z = (zc *)malloc(sizeof(zc) * n * n);
  zc *B = (zc *)malloc(sizeof(zc) * n * n);
  for (size_t i = 0; i < (size_t)n * n; ++i) { U0z[i] = U0[i]; V0z[i] = V0[i]; }
  zgemm('t', 'n', n, n, n, 1.0, Ubar, n, U0z, n, 0.0, B, n);
  zgemm('t', 'n', n, n, n, 1.0, Vbar, n, V0z, n, 1.0, B, n);
  zc d = zdet(n, B);
  free(U0z); free(V0z); free(B);
  return csqrt(d);
}

/* ROKAPPAphiomega (PNAMP_Proj.f90:459-551). */
int or_rokappa(int n, const zc *A, const double *U0, const double *V0, const double *U1,
               const double *V1, zc *rho, zc *k10, zc *k01) {
  size_t nn = (size_t)n * n;
  zc *Ac = malloc(sizeof(zc) * nn), *U0z = malloc(sizeof(zc) * nn), *V0z = malloc(sizeof(zc) * nn);
  zc *U1z = malloc(sizeof(zc) * nn), *V1z = malloc(sizeof(zc) * nn);
  zc *a1 = calloc(nn, sizeof(zc)), *a2 = calloc(nn, sizeof(zc)), *a3 = calloc(nn, sizeof(zc)),
     *a4 = calloc(nn, sizeof(zc)), *T = calloc(nn, sizeof(zc));
  for (size_t i = 0; i < nn; ++i) {
    Ac[i] = conj(A[i]); U0z[i] = U0[i]; V0z[i] = V0[i]; U1z[i] = U1[i]; V1z[i] = V1[i];
  }
  zgemm('t', 'n', n, n, n, 1.0, V0z, n, A, n, 0.0, a1, n);
  zgemm('n', 'n', n, n, n, 1.0, a1, n, V1z, n, 0.0, a2, n);
  zgemm('t', 'n', n, n, n, 1.0, U0z, n, Ac, n, 0.0, a3, n);
  zgemm('n', 'n', n, n, n, 1.0, a3, n, U1z, n, 0.0, a4, n);
  for (size_t i = 0; i < nn; ++i) T[i] = a2[i] + a4[i];
  int info = zinverse(n, T);
  zgemm('n', 'n', n, n, n, 1.0, A, n, V1z, n, 0.0, a1, n);
  zgemm('n', 'n', n, n, n, -1.0, Ac, n, U1z, n, 0.0, a2, n);
  zgemm('n', 't', n, n, n, 1.0, T, n, V0z, n, 0.0, a3, n);
  zgemm('n', 't', n, n, n, 1.0, T, n, U0z, n, 0.0, a4, n);
  zgemm('n', 'n', n, n, n, 1.0, a1, n, a3, n, 0.0, rho, n);
  zgemm('n', 'n', n, n, n, 1.0, a1, n, a4, n, 0.0, k10, n);
  zgemm('n', 'n', n, n, n, 1.0, a2, n, a3, n, 0.0, k01, n);
  free(Ac); free(U0z); free(V0z); free(U1z); free(V1z);
  free(a1); free(a2); free(a3); free(a4); free(T);
  return info;
}

/* ZN_0 (PNAMP_ME1B.f90:218-237). */
static void zn0(int n, const zc *rho, zc *zp, zc *zn) {
  *zp = 0; *zn = 0;
  for (int i = 0; i < n / 2; ++i) { *zp += rho[IX(i, i, n)]; *zn += rho[IX(i + n / 2, i + n / 2, n)]; }
}

/* HFB_Energy_COMPLEX (PNAMP_Energy_OpenMP.f90:5-100), Method II: returns
 * E2 + Ekin in *ehfb and Ekin in *ekin. */
void or_energy(const or_cfg *c, const zc *ro, const zc *k10, const zc *k01, zc *ekin, zc *ehfb) {
  int n = c->n;
#define RO(a, b) ro[IX((a) - 1, (b) - 1, n)]
#define K10(a, b) k10[IX((a) - 1, (b) - 1, n)]
#define K01(a, b) k01[IX((a) - 1, (b) - 1, n)]
  zc e = 0;
  for (int64_t q = 0; q < c->nme; ++q) {
    int i = c->ka[q], j = c->kb[q], k = c->kc[q], l = c->kd[q];
    double v = c->me[q];
    zc r_ijkl = RO(k, i) * RO(l, j) - RO(l, i) * RO(k, j) + K01(i, j) * K10(k, l);
    zc r_jikl = RO(k, j) * RO(l, i) - RO(l, j) * RO(k, i) + K01(j, i) * K10(k, l);
    zc r_jilk = RO(l, j) * RO(k, i) - RO(k, j) * RO(l, i) + K01(j, i) * K10(l, k);
    zc r_ijlk = RO(l, i) * RO(k, j) - RO(k, i) * RO(l, j) + K01(i, j) * K10(l, k);
    zc r_klij = RO(i, k) * RO(j, l) - RO(j, k) * RO(i, l) + K01(k, l) * K10(i, j);
    zc r_klji = RO(j, k) * RO(i, l) - RO(i, k) * RO(j, l) + K01(k, l) * K10(j, i);
    zc r_lkij = RO(i, l) * RO(j, k) - RO(j, l) * RO(i, k) + K01(l, k) * K10(i, j);
    zc r_lkji = RO(j, l) * RO(i, k) - RO(i, l) * RO(j, k) + K01(l, k) * K10(j, i);
    double d = (k + l == i + j) ? 0.0 : 1.0;
    e += 0.25 * v * (r_ijkl + d * r_klij - r_jikl - d * r_klji - r_ijlk - d * r_lkij + r_jilk + d * r_lkji);
  }
#undef RO
#undef K10
#undef K01
  zc ek = 0;
  for (int ia = 0; ia < n; ++ia)
    for (int ic = 0; ic < n; ++ic) ek += ro[IX(ic, ia, n)] * c->t1b[IX(ia, ic, n)];
  *ekin = ek;
  *ehfb = e + ek;
}

/* HFB_FIELD_COMPLEX + HFB_ENER_comp (VAPHFB_HFB_Field.f90:4-123,
 * PNAMP_HFB_Energy.f90:1-35), Method I, all mixing flags on. */
void or_energy_field(const or_cfg *c, const zc *ro, const zc *k10, const zc *k01, zc *ekin, zc *ehfb) {
  int n = c->n;
  size_t nn = (size_t)n * n;
  zc *G = calloc(nn, sizeof(zc)), *D10 = calloc(nn, sizeof(zc));
#define GA(a, b) G[IX((a) - 1, (b) - 1, n)]
#define DA(a, b) D10[IX((a) - 1, (b) - 1, n)]
#define RO(a, b) ro[IX((a) - 1, (b) - 1, n)]
#define K10(a, b) k10[IX((a) - 1, (b) - 1, n)]
  for (int64_t q = 0; q < c->nme; ++q) {
    int ia = c->ka[q], ib = c->kb[q], ic = c->kc[q], id = c->kd[q];
    double v = c->me[q];
    GA(ia, ic) += v * RO(id, ib); GA(ib, ic) -= v * RO(id, ia);
    GA(ia, id) -= v * RO(ic, ib); GA(ib, id) += v * RO(ic, ia);
    if (ia + ib != ic + id) {
      GA(ic, ia) += v * RO(ib, id); GA(id, ia) -= v * RO(ib, ic);
      GA(ic, ib) -= v * RO(ia, id); GA(id, ib) += v * RO(ia, ic);
    }
    DA(ia, ib) += 0.5 * v * K10(ic, id) * 2; DA(ib, ia) -= 0.5 * v * K10(ic, id) * 2;
    if (ia + ib != ic + id) {
      DA(ic, id) += 0.5 * v * K10(ia, ib) * 2; DA(id, ic) -= 0.5 * v * K10(ia, ib) * 2;
    }
  }
#undef GA
#undef DA
#undef RO
#undef K10
  zc ek = 0, ehf = 0, epa = 0;
  for (int i = 0; i < n; ++i)
    for (int l = 0; l < n; ++l) {
      ek += c->t1b[IX(i, l, n)] * ro[IX(l, i, n)];
      ehf += 0.5 * G[IX(i, l, n)] * ro[IX(l, i, n)];
      epa += 0.5 * D10[IX(i, l, n)] * k01[IX(i, l, n)];
    }
  *ekin = ek;
  *ehfb = ek + ehf + epa;
  free(G); free(D10);
}

/* One-body angular momentum matrices <a|J_k|b> (k = x, y, z) and their
 * squares, the cME1B%ZJ*ME arrays the reference builds in SingleParticleME
 * (Codes/PNAMP/PNAMP_ME1B.f90:138-216): non-zero inside one orbit. */
static void jmats(const or_cfg *c, zc *J, zc *J2) {
  int n = c->n;
  size_t nn = (size_t)n * n;
  zc *Jp = calloc(nn, sizeof(zc));
  for (size_t i = 0; i < 3 * nn; ++i) J[i] = 0;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      if (c->t[a] != c->t[b] || c->nq[a] != c->nq[b] || c->l[a] != c->l[b] || c->twoj[a] != c->twoj[b]) continue;
      double j = c->twoj[a] / 2.0, ma = c->twom[a] / 2.0, mb = c->twom[b] / 2.0;
      if (fabs(ma - mb - 1.0) < 1e-9) Jp[IX(a, b, n)] = sqrt(j * (j + 1) - mb * (mb + 1));
      if (a == b) J[2 * nn + IX(a, a, n)] = ma;
    }
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      zc p = Jp[IX(a, b, n)], m = Jp[IX(b, a, n)];
      J[IX(a, b, n)] = 0.5 * (p + m);
      J[nn + IX(a, b, n)] = (p - m) / (2.0 * I);
    }
  for (int k = 0; k < 3; ++k) zgemm('n', 'n', n, n, n, 1.0, J + k * nn, n, J + k * nn, n, 0.0, J2 + k * nn, n);
  free(Jp);
}

/* ZJ2_0 (Codes/PNAMP/PNAMP_ME2B.f90:1-152): <J^2> = sum_k Tr(J_k^2 rho)
 * + Tr(J_k rho)^2 - Tr(J_k rho J_k rho) + Tr(J_k kappa10 J_k^T kappa01^T). */
zc or_j2(const or_cfg *c, const zc *J, const zc *J2, const zc *ro, const zc *k10, const zc *k01) {
  int n = c->n;
  size_t nn = (size_t)n * n;
  zc *a1 = malloc(sizeof(zc) * nn), *a2 = malloc(sizeof(zc) * nn), *a3 = malloc(sizeof(zc) * nn),
     *a4 = malloc(sizeof(zc) * nn), *a5 = malloc(sizeof(zc) * nn), *a6 = malloc(sizeof(zc) * nn);
  zc tot = 0;
  for (int k = 0; k < 3; ++k) {
    zgemm('n', 'n', n, n, n, 1.0, J2 + k * nn, n, ro, n, 0.0, a1, n);
    zgemm('n', 'n', n, n, n, 1.0, J + k * nn, n, ro, n, 0.0, a2, n);
    zgemm('n', 'n', n, n, n, 1.0, J + k * nn, n, k10, n, 0.0, a3, n);
    zgemm('t', 't', n, n, n, 1.0, J + k * nn, n, k01, n, 0.0, a4, n);
    zgemm('n', 'n', n, n, n, 1.0, a2, n, a2, n, 0.0, a5, n);
    zgemm('n', 'n', n, n, n, 1.0, a3, n, a4, n, 0.0, a6, n);
    zc t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    for (int i = 0; i < n; ++i) {
      t1 += a1[IX(i, i, n)]; t2 += a2[IX(i, i, n)]; t3 += a5[IX(i, i, n)]; t4 += a6[IX(i, i, n)];
    }
    tot += t1 + t2 * t2 - t3 + t4;
  }
  free(a1); free(a2); free(a3); free(a4); free(a5); free(a6);
  return tot;
}

/* Per-gauge-point observables: norm, Z, N, Ekin, E+E0, J^2 (the ZRES slots 1,2,3,4,7,8). */
#define NOBS 6

/* One Euler point: the Fomenko gauge loop of Kernel4Triaxial
 * (PNAMP_Kernel3D.f90:245-387), returns ZRES(iang, :) for the five observables. */
/* Orbits of one species half (basis order, contiguous runs of equal n, l, j;
 * both halves have the same orbits): first m-state, 2j, l.  Returns norb. */
static int species_orbits(const or_cfg *c, int *first, int *tj, int *l) {
  int h = c->n / 2, norb = 0;
  for (int i = 0; i < h;) {
    first[norb] = i;
    tj[norb] = c->twoj[i];
    l[norb] = c->l[i];
    i += c->twoj[i] + 1;
    ++norb;
  }
  return norb;
}

/* Per gauge point, E2 and one-body-density observables (ZQ2,
 * Codes/PNAMP/PNAMP_ME1B.f90:52-75, and ZRho1B_Angles,
 * Codes/PNAMP/PNAMP_Rho1B.f90:2-36):
 *   q[s][mu] = sum_{i in s, j} Q_mu(i, j) rho(j, i)      (trace of Q_mu rho over species s)
 *   d[s][a][b] = sum_m rho(k_a(m), k_b(m))              (orbits a, b of species s with equal l, j)
 * q: [2][5], d: [2][norb][norb]. */
static void e2_point(const or_cfg *c, const zc *ro, zc *q, zc *d) {
  int n = c->n, h = n / 2;
  for (int sp = 0; sp < 2; ++sp)
    for (int mu = 0; mu < 5; ++mu) {
      zc acc = 0;
      const double *Q = c->q2 + (size_t)mu * n * n;
      for (int i = sp * h; i < sp * h + h; ++i)
        for (int j = 0; j < n; ++j) acc += Q[IX(i, j, n)] * ro[IX(j, i, n)];
      q[sp * 5 + mu] = acc;
    }
  int first[64], tj[64], l[64];
  int norb = species_orbits(c, first, tj, l);
  for (int sp = 0; sp < 2; ++sp)
    for (int a = 0; a < norb; ++a)
      for (int b = 0; b < norb; ++b) {
        zc acc = 0;
        if (tj[a] == tj[b] && l[a] == l[b])
          for (int m = 0; m <= tj[a]; ++m) acc += ro[IX(sp * h + first[a] + m, sp * h + first[b] + m, n)];
        d[((size_t)sp * norb + a) * norb + b] = acc;
      }
}

static int euler_point(const or_cfg *c, double alp, double bet, double gam, const double *U0,
                       const double *V0, const double *U1, const double *V1, const zc *Zc0,
                       double det0, double det1, zc *zres, zc *ze2, zc *zrho) {
  int n = c->n;
  size_t nn = (size_t)n * n;
  zc *R = malloc(sizeof(zc) * nn), *Ub = malloc(sizeof(zc) * nn), *Vb = malloc(sizeof(zc) * nn),
     *A = malloc(sizeof(zc) * nn), *ro = malloc(sizeof(zc) * nn), *k10 = malloc(sizeof(zc) * nn),
     *k01 = malloc(sizeof(zc) * nn), *J = malloc(sizeof(zc) * 3 * nn), *J2 = malloc(sizeof(zc) * 3 * nn);
  jmats(c, J, J2);
  or_rotation_mat(c, alp, bet, gam, R);
  double aaa = (double)n * (n + 1) / 2.0;
  double snpfaf = ((long)aaa % 2) ? -1.0 : 1.0;
  double scale = pow(10.0, (double)(-(n * c->iscale) / 2));
  zc sum[NOBS] = {0};
  int nf = c->nphi;
  int info = 0;
  int first[64], tjo[64], lo[64];
  int norb = species_orbits(c, first, tjo, lo);
  zc qs[10], *ds = malloc(sizeof(zc) * 2 * norb * norb);
  zc sq[10] = {0}, *sd = calloc(2 * norb * norb, sizeof(zc));
  for (int lp = 0; lp < nf; ++lp) {
    double phip = (nf == 1) ? 0.0 : (2 * M_PI / nf) * lp;
    zc zpp = cexp(-I * phip * c->zval);
    for (int ln = 0; ln < nf; ++ln) {
      double phin = (nf == 1) ? 0.0 : (2 * M_PI / nf) * ln;
      zc zpn = cexp(-I * phin * c->nval);
      or_uvphiomega(n, phip, phin, R, U1, V1, Ub, Vb, A);
      zc znorm;
      if (c->norm_method == 1) {
        znorm = or_onishi(n, U0, V0, Ub, Vb);
      } else {
        zc pf;
        int e = or_pfaf_over(n, Zc0, Ub, Vb, c->cscale, &pf);
        if (e && !info) info = e;
        znorm = snpfaf * (det0 * scale) * (det1 * scale) * pf * zpp * zpn;
      }
      int e = or_rokappa(n, A, U0, V0, U1, V1, ro, k10, k01);
      if (e && !info) info = e;
      zc zp, zn, ek, eh;
      zn0(n, ro, &zp, &zn);
      if (c->energy_method == 1) or_energy_field(c, ro, k10, k01, &ek, &eh);
      else or_energy(c, ro, k10, k01, &ek, &eh);
      sum[0] += znorm;
      sum[1] += znorm * zp;
      sum[2] += znorm * zn;
      sum[3] += znorm * ek;
      sum[4] += znorm * (eh + c->e0);
      sum[5] += znorm * or_j2(c, J, J2, ro, k10, k01);
      if (ze2 && c->q2) {
        e2_point(c, ro, qs, ds);
        for (int k = 0; k < 10; ++k) sq[k] += znorm * qs[k];
        for (int k = 0; k < 2 * norb * norb; ++k) sd[k] += znorm * ds[k];
      }
    }
  }
  for (int io = 0; io < NOBS; ++io) zres[io] = sum[io] / (double)(nf * nf);
  if (ze2 && c->q2) {
    for (int k = 0; k < 10; ++k) ze2[k] = sq[k] / (double)(nf * nf);
    for (int k = 0; k < 2 * norb * norb; ++k) zrho[k] = sd[k] / (double)(nf * nf);
  }
  free(ds);
  free(sd);
  free(R); free(Ub); free(Vb); free(A); free(ro); free(k10); free(k01); free(J); free(J2);
  return info;
}

/* J-projection.  fold=0: plain  N^J_{K1K2} = wJ sum_O w D^{J*}_{K1K2}(O) ZRES(O),
 * D^{J*} = e^{i K1 alp} d^J_{K1K2}(bet) e^{i K2 gam} (PNAMP_Kernel3D.f90:8), all J.
 * fold=1: the reference text PNAMP_Kernel3D.f90:415-526 (even J, four-term mirror fold). */
void or_project(const or_cfg *c, const zc *zres, zc *kern) {
  int J2m = 2 * c->jmax + 1;
  int na = c->nalp, nb = c->nbet, ng = c->ngam;
  size_t per_io = (size_t)(c->jmax + 1) * J2m * J2m;
  memset(kern, 0, sizeof(zc) * per_io * NOBS);
  int onept = (na * ng == 1);
  for (int J = 0; J <= c->jmax; ++J) {
    if (c->fold == 1 && (J % 2)) continue;
    double wJ = onept ? (2.0 * J + 1.0) / 2.0
                      : (c->fold == 1 ? 2.0 : 1.0) * (2.0 * J + 1.0) / (8.0 * M_PI * M_PI);
    for (int K1 = -J; K1 <= J; ++K1)
      for (int K2 = -J; K2 <= J; ++K2) {
        zc acc[NOBS] = {0};
        for (int ib = 0; ib < nb; ++ib) {
          double bet = (nb == 1) ? 0.0 : c->bet[ib];
          double wb = (nb == 1) ? 1.0 : c->wbet[ib];
          double dp = or_wigner_d(2 * J, 2 * K1, 2 * K2, bet);
          double dm = or_wigner_d(2 * J, 2 * K1, 2 * K2, M_PI - bet);
          for (int ia = 0; ia < na; ++ia)
            for (int ig = 0; ig < ng; ++ig) {
              double alp = c->alp[ia], gam = c->gam[ig];
              double w = c->walp[ia] * wb * c->wgam[ig];
              int i1 = (ib * na + ia) * ng + ig;
              if (c->fold == 0) {
                zc D = cexp(I * K1 * alp) * dp * cexp(I * K2 * gam);
                for (int io = 0; io < NOBS; ++io) acc[io] += wJ * D * w * zres[(size_t)i1 * NOBS + io];
              } else {
                int i2 = (ib * na + ia) * ng + (ng - 1 - ig);
                int i3 = (ib * na + (na - 1 - ia)) * ng + ig;
                int i4 = (ib * na + (na - 1 - ia)) * ng + (ng - 1 - ig);
                zc zap = cexp(I * K1 * alp), zam = cexp(-I * K1 * alp), zgp = cexp(I * K2 * gam);
                zc za1 = zap * zgp * dp, za2 = zam * zgp * dp, za3 = zap * zgp * dm, za4 = zam * zgp * dm;
                double d0 = (K1 == 0 ? 2.0 : 1.0) * (K2 == 0 ? 2.0 : 1.0);
                for (int io = 0; io < NOBS; ++io) {
                  zc zo1 = zres[(size_t)i1 * NOBS + io], zo2 = conj(zres[(size_t)i2 * NOBS + io]);
                  zc zo3 = conj(zres[(size_t)i3 * NOBS + io]), zo4 = zres[(size_t)i4 * NOBS + io];
                  acc[io] += wJ * (za1 * zo1 + za2 * zo2 + za3 * zo3 + za4 * zo4) * w / d0;
                }
              }
            }
        }
        for (int io = 0; io < NOBS; ++io)
          kern[(size_t)io * per_io + ((size_t)J * J2m + (K1 + c->jmax)) * J2m + (K2 + c->jmax)] = acc[io];
      }
  }
}

/* Kernel4Triaxial for one pair: zres [NE][NOBS] (may be NULL),
 * kern [NOBS][jmax+1][2jmax+1][2jmax+1] (may be NULL).  Threads over Euler
 * points (the reference spreads pairs over MPI ranks; the per-point work is
 * the same).  Returns 0, or the first LAPACK-style INFO. */
int or_kernel_pair(const or_cfg *c, const double *U0, const double *V0, const double *U1,
                   const double *V1, zc *zres_out, zc *kern) {
  int n = c->n;
  int NE = c->nalp * c->nbet * c->ngam;
  zc *Zc0 = malloc(sizeof(zc) * n * n), *Zc1 = malloc(sizeof(zc) * n * n);
  double det0 = 1, det1 = 1;
  int info = or_zbareover(n, U0, V0, Zc0, &det0);
  if (!info) info = or_zbareover(n, U1, V1, Zc1, &det1);
  if (info) { free(Zc0); free(Zc1); return info; }
  zc *zres = zres_out ? zres_out : malloc(sizeof(zc) * NE * NOBS);
  int err = 0;
#ifdef _OPENMP
  int nt = c->nthreads > 0 ? c->nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(| : err)
#endif
  for (int iang = 0; iang < NE; ++iang) {
    int ig = iang % c->ngam, ia = (iang / c->ngam) % c->nalp, ib = iang / (c->ngam * c->nalp);
    double bet = (c->nbet == 1) ? 0.0 : c->bet[ib];
    err |= euler_point(c, c->alp[ia], bet, c->gam[ig], U0, V0, U1, V1, Zc0, det0, det1,
                       zres + (size_t)iang * NOBS, NULL, NULL);
  }
  if (kern) or_project(c, zres, kern);
  if (!zres_out) free(zres);
  free(Zc0); free(Zc1);
  return err ? 1 : 0;
}

/* Reduced E2 matrix elements and the J = 0 one-body transition density of
 * one pair (the iE2 = 1 part of Kernel4Triaxial).
 *
 * ze2 [NE][2][5]: per Euler point, norm-weighted gauge sums of tr_s(Q_{2mu} rho)
 * (ZRES slots 9-13 of the reference, both species); zrho [NE][2][norb][norb]
 * the same for the orbit densities (ZOBDJ0).
 *
 * qred [2][5][jmax+1][2jmax+5][2jmax+1] (species, J1 - J2 + 2, J2, K1 + jmax + 2,
 * K2 + jmax), the reduced matrix elements <J1 K1 q0 || Q2 || J2 K2 q1> of
 * PNAMP_Kernel3D.f90:590-710:
 *   wJ(J2) (2 J1 + 1) (-1)^(J1 - K1) sum_mu (J1 2 J2; -K1 mu K1-mu)
 *     sum_O w D^{J2*}_{K1-mu, K2}(O) ze2_mu(O),   |K1 - mu| <= J2,
 * with D^{J*}_{K'K} = e^{i K' alp} d^J_{K'K}(bet) e^{i K gam} and wJ as in
 * or_project (plain fold).  By 3j orthogonality this is the Wigner-Eckart
 * reduced element: <Phi0| P^{J1}_{K1M1} Q_{2mu} P^{J2}_{M2K2} |Phi1> =
 * (-1)^(J1-M1) (J1 2 J2; -M1 mu M2) <J1 K1 || Q2 || J2 K2>.
 * rho1b [2][norb][norb]: ZRho1B_Integral (Codes/PNAMP/PNAMP_Rho1B.f90:39-62),
 * wJ(0) sum_O w zrho(O) / sqrt(2 j_a + 1) (the 1/sqrt(2j+1) of
 * ZRho1B_Angles, :29-30). */
static void or_project_e2(const or_cfg *c, const zc *ze2, const zc *zrho, zc *qred, zc *rho1b) {
  int na = c->nalp, nb = c->nbet, ng = c->ngam, jm = c->jmax;
  int K1d = 2 * jm + 5, K2d = 2 * jm + 1;
  int onept = (na * ng == 1);
  memset(qred, 0, sizeof(zc) * 2 * 5 * (jm + 1) * K1d * K2d);
  for (int J2 = 0; J2 <= jm; ++J2) {
    double wJ = onept ? (2.0 * J2 + 1.0) / 2.0 : (2.0 * J2 + 1.0) / (8.0 * M_PI * M_PI);
    for (int dJ = -2; dJ <= 2; ++dJ) {
      int J1 = J2 + dJ;
      if (J1 < 0) continue;
      for (int K1 = -J1; K1 <= J1; ++K1)
        for (int K2 = -J2; K2 <= J2; ++K2) {
          zc acc[2] = {0, 0};
          for (int mu = -2; mu <= 2; ++mu) {
            int Kp = K1 - mu;
            if (Kp < -J2 || Kp > J2) continue;
            double tj = or_threej(2 * J1, 4, 2 * J2, -2 * K1, 2 * mu, 2 * Kp);
            if (tj == 0.0) continue;
            for (int ib = 0; ib < nb; ++ib) {
              double bet = (nb == 1) ? 0.0 : c->bet[ib], wb = (nb == 1) ? 1.0 : c->wbet[ib];
              double dp = or_wigner_d(2 * J2, 2 * Kp, 2 * K2, bet);
              for (int ia = 0; ia < na; ++ia)
                for (int ig = 0; ig < ng; ++ig) {
                  int i1 = (ib * na + ia) * ng + ig;
                  zc D = cexp(I * Kp * c->alp[ia]) * dp * cexp(I * K2 * c->gam[ig]);
                  double w = c->walp[ia] * wb * c->wgam[ig];
                  for (int sp = 0; sp < 2; ++sp) acc[sp] += tj * w * D * ze2[(size_t)i1 * 10 + sp * 5 + (mu + 2)];
                }
            }
          }
          double ph = ((J1 - K1) & 1) ? -1.0 : 1.0;
          for (int sp = 0; sp < 2; ++sp)
            qred[((((size_t)sp * 5 + (dJ + 2)) * (jm + 1) + J2) * K1d + (K1 + jm + 2)) * K2d + (K2 + jm)] =
                wJ * (2 * J1 + 1) * ph * acc[sp];
        }
    }
  }
  int first[64], tj[64], lo[64];
  int norb = species_orbits(c, first, tj, lo);
  double w0 = onept ? 0.5 : 1.0 / (8.0 * M_PI * M_PI);
  for (int k = 0; k < 2 * norb * norb; ++k) {
    zc acc = 0;
    for (int ib = 0; ib < nb; ++ib)
      for (int ia = 0; ia < na; ++ia)
        for (int ig = 0; ig < ng; ++ig) {
          double wb = (nb == 1) ? 1.0 : c->wbet[ib];
          acc += c->walp[ia] * wb * c->wgam[ig] * zrho[(size_t)((ib * na + ia) * ng + ig) * 2 * norb * norb + k];
        }
    int a = (k / norb) % norb;
    rho1b[k] = w0 * acc / sqrt(tj[a] + 1.0);
  }
}

int or_norb(const or_cfg *c) {
  int first[64], tj[64], l[64];
  return species_orbits(c, first, tj, l);
}

/* Kernel4Triaxial with iE2 = 1 for one pair: the outputs of or_kernel_pair
 * plus ze2, zrho, qred, rho1b (layouts at or_project_e2; any may be NULL
 * except that qred / rho1b need c->q2). */
int or_kernel_pair_e2(const or_cfg *c, const double *U0, const double *V0, const double *U1, const double *V1,
                      zc *zres_out, zc *kern, zc *ze2_out, zc *zrho_out, zc *qred, zc *rho1b) {
  if (!c->q2) return -1;
  int n = c->n;
  int NE = c->nalp * c->nbet * c->ngam;
  int norb = or_norb(c);
  zc *Zc0 = malloc(sizeof(zc) * n * n), *Zc1 = malloc(sizeof(zc) * n * n);
  double det0 = 1, det1 = 1;
  int info = or_zbareover(n, U0, V0, Zc0, &det0);
  if (!info) info = or_zbareover(n, U1, V1, Zc1, &det1);
  if (info) { free(Zc0); free(Zc1); return info; }
  zc *zres = zres_out ? zres_out : malloc(sizeof(zc) * NE * NOBS);
  zc *ze2 = ze2_out ? ze2_out : malloc(sizeof(zc) * NE * 10);
  zc *zrho = zrho_out ? zrho_out : malloc(sizeof(zc) * NE * 2 * norb * norb);
  int err = 0;
#ifdef _OPENMP
  int nt = c->nthreads > 0 ? c->nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(| : err)
#endif
  for (int iang = 0; iang < NE; ++iang) {
    int ig = iang % c->ngam, ia = (iang / c->ngam) % c->nalp, ib = iang / (c->ngam * c->nalp);
    double bet = (c->nbet == 1) ? 0.0 : c->bet[ib];
    err |= euler_point(c, c->alp[ia], bet, c->gam[ig], U0, V0, U1, V1, Zc0, det0, det1,
                       zres + (size_t)iang * NOBS, ze2 + (size_t)iang * 10, zrho + (size_t)iang * 2 * norb * norb);
  }
  if (kern) or_project(c, zres, kern);
  if (qred && rho1b) or_project_e2(c, ze2, zrho, qred, rho1b);
  if (!zres_out) free(zres);
  if (!ze2_out) free(ze2);
  if (!zrho_out) free(zrho);
  free(Zc0); free(Zc1);
  return err ? 1 : 0;
}

/* Per-point primitives for the tests: one gauge point (alp,bet,gam,phip,phin). */
int or_point(const or_cfg *c, double alp, double bet, double gam, double phip, double phin,
             const double *U0, const double *V0, const double *U1, const double *V1, zc *out_norm,
             zc *rho, zc *k10, zc *k01, zc *ekin, zc *ehfb, zc *ekin_f, zc *ehfb_f) {
  int n = c->n;
  size_t nn = (size_t)n * n;
  zc *R = malloc(sizeof(zc) * nn), *Ub = malloc(sizeof(zc) * nn), *Vb = malloc(sizeof(zc) * nn),
     *A = malloc(sizeof(zc) * nn), *Zc0 = malloc(sizeof(zc) * nn), *Zc1 = malloc(sizeof(zc) * nn);
  double det0, det1;
  int info = or_zbareover(n, U0, V0, Zc0, &det0);
  if (!info) info = or_zbareover(n, U1, V1, Zc1, &det1);
  if (!info) {
    or_rotation_mat(c, alp, bet, gam, R);
    or_uvphiomega(n, phip, phin, R, U1, V1, Ub, Vb, A);
    double aaa = (double)n * (n + 1) / 2.0;
    double snpfaf = ((long)aaa % 2) ? -1.0 : 1.0;
    double scale = pow(10.0, (double)(-(n * c->iscale) / 2));
    zc pf;
    info = or_pfaf_over(n, Zc0, Ub, Vb, c->cscale, &pf);
    out_norm[0] = snpfaf * det0 * scale * det1 * scale * pf * cexp(-I * phip * c->zval) * cexp(-I * phin * c->nval);
    out_norm[1] = or_onishi(n, U0, V0, Ub, Vb);
    out_norm[2] = det0;
    out_norm[3] = det1;
    int e = or_rokappa(n, A, U0, V0, U1, V1, rho, k10, k01);
    if (e && !info) info = e;
    or_energy(c, rho, k10, k01, ekin, ehfb);
    or_energy_field(c, rho, k10, k01, ekin_f, ehfb_f);
  }
  free(R); free(Ub); free(Vb); free(A); free(Zc0); free(Zc1);
  return info;
}

/* Exposed helpers. */
void or_pfaffian(int N, zc *S) { /* S in place; result returned in S[0] */
  zc p = or_zpfaffianf(S, N, N);
  S[0] = p;
}

/* HNDIAG (Codes/GCM/GCM_HNDiag.f90:2-103) with sdiag restated as Jacobi.
 * NN, HH real symmetric dim x dim (column-major).  On return: EN (descending
 * norm eigenvalues, dim), E (ascending energies, first *m), F (dim x *m
 * collective amplitudes FF = RR GG), returns IFL. */
int or_hndiag(int dim, const double *NN, const double *HH, double eps, int nos, double *EN,
              int *m_out, double *E, double *F) {
  double *DD = malloc(sizeof(double) * dim * dim), *vec = malloc(sizeof(double) * dim * dim);
  for (int i = 0; i < dim * dim; ++i) DD[i] = -NN[i];
  double *w = malloc(sizeof(double) * dim);
  jacobi_eig(dim, DD, w, vec);
  for (int i = 0; i < dim; ++i) EN[i] = -w[i];
  int ik = 0;
  for (ik = 0; ik < dim; ++ik)
    if (EN[ik] / EN[0] < eps) break;
  int M = ik;
  if (nos > 0 && nos < M) M = nos;
  *m_out = M;
  if (M <= 0) { free(DD); free(vec); free(w); return 1; }
  double *RR = malloc(sizeof(double) * dim * M), *GG = malloc(sizeof(double) * M * M), *gv = malloc(sizeof(double) * M * M);
  for (int k = 0; k < M; ++k) {
    double s = 1.0 / sqrt(EN[k]);
    for (int i = 0; i < dim; ++i) RR[IX(i, k, dim)] = vec[IX(i, k, dim)] * s;
  }
  for (int k = 0; k < M; ++k)
    for (int j = 0; j < M; ++j) {
      double s = 0;
      for (int i = 0; i < dim; ++i)
        for (int l = 0; l < dim; ++l) s += RR[IX(i, j, dim)] * HH[IX(i, l, dim)] * RR[IX(l, k, dim)];
      GG[IX(j, k, M)] = s;
    }
  for (int k = 0; k < M; ++k) /* symmetrise from the computed triangle like HNDIAG (J>=K) */
    for (int j = k + 1; j < M; ++j) GG[IX(k, j, M)] = GG[IX(j, k, M)];
  jacobi_eig(M, GG, E, gv);
  if (F)
    for (int k = 0; k < M; ++k)
      for (int i = 0; i < dim; ++i) {
        double s = 0;
        for (int l = 0; l < M; ++l) s += RR[IX(i, l, dim)] * gv[IX(l, k, M)];
        F[IX(i, k, dim)] = s;
      }
  free(DD); free(vec); free(w); free(RR); free(GG); free(gv);
  return 0;
}
