// pnamp_dev.h -- complex FP64 helpers and wavefront reductions shared by the
// HIP sources of the PNAMP library (gfx950, 64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>

namespace pnamp_dev {

__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ void cfms(double2 &acc, double2 a, double2 b) {  // acc -= a * b
  acc.x = fma(-a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(-a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
__device__ __forceinline__ void rfma(double2 &acc, double a, double2 b) {  // acc += a * b, a real
  acc.x = fma(a, b.x, acc.x);
  acc.y = fma(a, b.y, acc.y);
}
__device__ __forceinline__ double2 crecip(double2 a) {
  double d = a.x * a.x + a.y * a.y;
  return make_double2(a.x / d, -a.y / d);
}
__device__ __forceinline__ double2 crecip1(double2 a) {  // one division
  const double r = 1.0 / (a.x * a.x + a.y * a.y);
  return make_double2(a.x * r, -a.y * r);
}
// Read-only view in the constant address space: wave-uniform reads through it
// become scalar loads (a generic pointer that stores in the kernel might
// alias is read with per-lane vector loads)
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *ro_const(const T *q) {
  return (const __attribute__((address_space(4))) T *)q;
}
__device__ __forceinline__ double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }
__device__ __forceinline__ double2 cexpi(double t) {
  double s, c;
  sincos(t, &s, &c);
  return make_double2(c, s);
}
__device__ __forceinline__ double2 shfl_xor2(double2 v, int m) {
  return make_double2(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
}

// Wavefront argmax: the largest v, ties to the smaller key k, on every lane.
__device__ __forceinline__ void wave_best(double &v, int &k) {
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int ok = __shfl_xor(k, off);
    if (ov > v || (ov == v && ok < k)) { v = ov; k = ok; }
  }
}

}  // namespace pnamp_dev
