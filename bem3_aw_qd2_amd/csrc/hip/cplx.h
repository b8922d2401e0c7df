// cplx.h -- minimal double-precision complex arithmetic for device code.
#pragma once
#include <hip/hip_runtime.h>

struct cd {
  double x, y;
};

__device__ __forceinline__ cd mk(double x, double y) { return cd{x, y}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.x, -a.y}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
  return cd{fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ cd operator*(cd a, double s) { return cd{a.x * s, a.y * s}; }
__device__ __forceinline__ cd operator*(double s, cd a) { return cd{a.x * s, a.y * s}; }
__device__ __forceinline__ cd& operator+=(cd& a, cd b) { a.x += b.x; a.y += b.y; return a; }
__device__ __forceinline__ cd& operator-=(cd& a, cd b) { a.x -= b.x; a.y -= b.y; return a; }
__device__ __forceinline__ cd& operator*=(cd& a, cd b) { a = a * b; return a; }
__device__ __forceinline__ cd& operator*=(cd& a, double s) { a.x *= s; a.y *= s; return a; }
// 1/x and 1/sqrt(x) for finite x > 0 (no special operands): the hardware
// estimate refined by two Newton-Raphson steps, each of which squares the
// relative error, to within an ulp or two of the rounded result
__device__ __forceinline__ double rcp_pos(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double rsqrt_pos(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  return fma(y, e, y);
}
// 1/a for a != 0 by rcp_pos of |a|^2
__device__ __forceinline__ cd crecip_nz(cd a) {
  double d = rcp_pos(fma(a.x, a.x, a.y * a.y));
  return cd{a.x * d, -a.y * d};
}
__device__ __forceinline__ cd crecip(cd a) {
  double d = 1.0 / fma(a.x, a.x, a.y * a.y);
  return cd{a.x * d, -a.y * d};
}
__device__ __forceinline__ cd operator/(cd a, cd b) { return a * crecip(b); }
__device__ __forceinline__ cd operator/(cd a, double s) { double r = 1.0 / s; return cd{a.x * r, a.y * r}; }
__device__ __forceinline__ cd cis(double t) {
  double s, c;
  sincos(t, &s, &c);
  return cd{c, s};
}
// multiply by i
__device__ __forceinline__ cd mul_i(cd a) { return cd{-a.y, a.x}; }
__device__ __forceinline__ cd conj(cd a) { return cd{a.x, -a.y}; }
