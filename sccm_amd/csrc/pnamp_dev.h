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
// 1 / d for a positive normal d: the hardware reciprocal estimate (v_rcp_f64)
// refined by two Newton-Raphson steps r <- r + r (1 - d r), each of which
// about doubles the number of correct bits (a few instructions instead of the
// scaled division sequence; within an ulp or two of the rounded quotient)
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double2 crecip_nr(double2 a) {
  const double r = rcp_nr(a.x * a.x + a.y * a.y);
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

// Maximum over the wavefront by DPP (quad exchanges, row mirrors, row
// broadcasts: lane 63 ends with the maximum), returned wave-uniform; lanes
// without a DPP source read 0, the identity of an unsigned maximum.
template <int CTRL, int RM>
__device__ __forceinline__ unsigned dpp_umax(unsigned v) {
  return max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, RM, 0xF, false));
}
__device__ __forceinline__ unsigned wave_umax(unsigned v) {
  v = dpp_umax<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_umax<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_umax<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_umax<0x140, 0xF>(v);  // row_mirror
  v = dpp_umax<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_umax<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}
template <int CTRL, int RM>
__device__ __forceinline__ unsigned long long dpp_umax64(unsigned long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, RM, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, RM, 0xF, false);
  const unsigned long long o = ((unsigned long long)hi << 32) | lo;
  return o > v ? o : v;
}
__device__ __forceinline__ unsigned long long wave_umax64(unsigned long long v) {
  v = dpp_umax64<0xB1, 0xF>(v);
  v = dpp_umax64<0x4E, 0xF>(v);
  v = dpp_umax64<0x141, 0xF>(v);
  v = dpp_umax64<0x140, 0xF>(v);
  v = dpp_umax64<0x142, 0xA>(v);
  v = dpp_umax64<0x143, 0xC>(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

// Pivot key: a non-negative magnitude as FP32 bits with the low KB bits
// replaced by (2^KB - 1 - idx), so one unsigned maximum over the wavefront
// picks the largest magnitude (to 2^-(23-KB) relative) and, among equal keys,
// the smallest index.  A NaN magnitude counts as zero.
template <int KB>
__device__ __forceinline__ unsigned pivot_key(double mag, int idx) {
  const float f = (mag >= 0.0) ? (float)mag : 0.0f;
  return (__float_as_uint(f) & ~((1u << KB) - 1u)) | ((1u << KB) - 1u - (unsigned)idx);
}

// Pivot key from the high word of a non-negative double (sign 0, 11 exponent
// and 20 mantissa bits: unsigned order is numeric order) with the low KB bits
// replaced by (2^KB - 1 - idx): the largest magnitude to 2^-(20-KB) relative,
// ties to the smallest index.  A NaN or infinite magnitude wins.
template <int KB>
__device__ __forceinline__ unsigned pivot_key_hi(double mag, int idx) {
  return ((unsigned)((unsigned long long)__double_as_longlong(mag) >> 32) & ~((1u << KB) - 1u)) |
         ((1u << KB) - 1u - (unsigned)idx);
}

// Maximum over each aligned group of eight lanes, on all eight (quad
// exchanges, then the half-row mirror).
__device__ __forceinline__ unsigned oct_umax(unsigned v) {
  v = dpp_umax<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_umax<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_umax<0x141, 0xF>(v);  // row_half_mirror
  return v;
}

// Both halves of a wavefront's value, lanes 0..31 (lo) and 32..63 (hi), on
// every lane: v_permlane32_swap exchanges the upper half of one operand with
// the lower half of the other.
__device__ __forceinline__ void halves(double v, double &lo, double &hi) {
  const long long b = __double_as_longlong(v);
  const auto l = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  const auto h = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  lo = __longlong_as_double(((long long)h[0] << 32) | (long long)l[0]);
  hi = __longlong_as_double(((long long)h[1] << 32) | (long long)l[1]);
}
__device__ __forceinline__ void halves2(double2 v, double2 &lo, double2 &hi) {
  halves(v.x, lo.x, hi.x);
  halves(v.y, lo.y, hi.y);
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
