// plain_gpu.h -- host interface of the plain-form mesh-point kernels
// (plain_gpu.hip), used by pnamp_gpu.hip.
//
// The plain form works on units = (pair, Euler point), numbered
// u = pair * n_euler + euler.  For a range of units it runs
//   k_plain_setup  per unit: the gauge-independent products of the rotation
//                  with the states (7 n x n complex matrices) -> work buffer;
//   k_plain_gauge  per (unit, gauge point): Pfaffian, T^-1, densities, energy,
//                  <J^2>, traces (and the E2 traces) -> work buffer;
//   k_plain_sum    per unit: the Fomenko sum over the gauge points -> ZRES.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

struct PlainArgs {
  int n, h, nalp, nbet, ngam, NE, nphi, npack;
  int zval, nval;
  double cscale, vac_scale, snpfaf, e0;
  const int *q0, *q1;              // [npairs] state indices (device)
  const int *pz, *pn;              // [npairs] valence numbers per pair, or NULL
  const double *U, *V, *Z, *det;   // states (device)
  const double *alp, *gam, *rtab;  // Euler angles, d-tables [nbet][npack]
  const int *pk_a, *pk_b;          // packed rotation entries -> local (a, b)
  const int *blk_off, *blk_sz, *blk_base, *twom;  // per species-local m-state
  const double *mtab;              // [3][h]: m, <m+1|J+|m> (0 at m = j), j(j+1)
  const double *t1b;               // one-body kinetic n x n, column-major
  const uint4 *lofs;               // energy-list terms: LDS offsets
  const double *lme;               // their matrix elements
  int nl, lbytes;                  // terms (multiple of 256); 1: byte offsets, 0: element offsets
  // E2 traces (NULL q2: none)
  const double *q2;                // [5][n*n]
  const int4 *orbp;                // [nob2] orbit pairs
  int nob2;
  // outputs, indexed by global unit
  double2 *zres;                   // [units][PNAMP_NOBS]
  int *status;                     // [units]
  double2 *ze2, *zrho;             // [units][10], [units][nob2] (E2 only)
};

// n supported by the plain kernels: even, 8 <= n <= 64 (size classes
// 16, 24, 32, 40, 48, 64; the Pfaffian's 64-bit pivot keys carry the
// row-major rank of an entry of the 2n x 2n matrix in 14 bits, and the
// Gauss-Jordan keeps its used columns in a 64-bit mask)
bool plain_supported(int n);
// work-buffer bytes per unit
size_t plain_unit_bytes(const PlainArgs &a);
// Runs units [u0, u0 + nu) on stream s with the work buffer (at least
// nu * plain_unit_bytes bytes, 256-byte aligned).
hipError_t plain_run(const PlainArgs &a, int u0, int nu, void *work, hipStream_t s);
// Dynamic LDS attributes of the kernels for this n (per device, before plain_run).
hipError_t plain_prepare(int n, bool e2);
