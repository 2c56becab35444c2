// Shared device/host helpers for the MI355X (gfx950) RCWA solve path.
// Complex numbers are stored as interleaved (re, im) doubles, the layout of
// std::complex<double> and of Eigen's complex matrices (rcwa.hpp:12-13), so
// host buffers cross the C ABI without repacking.  Matrices are column-major.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

typedef double2 zc;  // complex128: .x = re, .y = im

__host__ __device__ __forceinline__ zc zmake(double r, double i) { zc z; z.x = r; z.y = i; return z; }
__host__ __device__ __forceinline__ zc zadd(zc a, zc b) { return zmake(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ zc zsub(zc a, zc b) { return zmake(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ zc zmul(zc a, zc b) {
  return zmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ zc zconj(zc a) { return zmake(a.x, -a.y); }
__host__ __device__ __forceinline__ zc zneg(zc a) { return zmake(-a.x, -a.y); }
__host__ __device__ __forceinline__ zc zscale(zc a, double s) { return zmake(a.x * s, a.y * s); }
// a*conj(b)
__host__ __device__ __forceinline__ zc zmulc(zc a, zc b) {
  return zmake(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// acc + a*b
__host__ __device__ __forceinline__ zc zfma(zc a, zc b, zc acc) {
  return zmake(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
__host__ __device__ __forceinline__ double zabs2(zc a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ __forceinline__ double zabs1(zc a) { return fabs(a.x) + fabs(a.y); }
__host__ __device__ __forceinline__ double zabs(zc a) { return hypot(a.x, a.y); }
__host__ __device__ __forceinline__ zc zdiv(zc a, zc b) {
  // Smith's algorithm (1962): scale by the larger component of b.
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return zmake((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.x * r + b.y;
    return zmake((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__host__ __device__ __forceinline__ zc zinv(zc b) { return zdiv(zmake(1.0, 0.0), b); }

// Principal square root, written from the half-angle identities
//   sqrt(a+ib) = t + i b/(2t),  t = sqrt((|z|+a)/2)        for a >= 0
//   sqrt(a+ib) = |b|/(2t) + i sign(b) t,  t = sqrt((|z|-a)/2) for a < 0
// (W. Kahan, "Branch cuts for complex elementary functions", 1987).  The sign
// of a zero imaginary part selects the side of the cut, as in C99 csqrt.
__host__ __device__ __forceinline__ zc zsqrt(zc z) {
  const double a = z.x, b = z.y;
  if (a == 0.0 && b == 0.0) return zmake(0.0, b);
  const double r = hypot(a, b);
  if (a >= 0.0) {
    const double t = sqrt(0.5 * (r + a));
    return zmake(t, b / (2.0 * t));
  }
  const double t = sqrt(0.5 * (r - a));
  return zmake(fabs(b) / (2.0 * t), copysign(t, b));
}
// rcwa.hpp:24-28: an imaginary part of -0 is treated as +0.
__host__ __device__ __forceinline__ zc zusqrt(zc z) {
  if (z.y == 0.0) z.y = 0.0;
  return zsqrt(z);
}
__device__ __forceinline__ zc zexp(zc z) {
  double s, c;
  sincos(z.y, &s, &c);
  const double e = exp(z.x);
  return zmake(e * c, e * s);
}

#define RCWA_HIP_CHECK(expr)                                   \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) return rcwa_internal_hip_fail(_e, #expr, __FILE__, __LINE__); \
  } while (0)

int rcwa_internal_hip_fail(hipError_t e, const char* what, const char* file, int line);

// Strided batch descriptor for matrices: element (i, j) of batch b is at
// base + b*stride + i + j*ld.
struct ZMat {
  zc* p;
  long long ld;
  long long stride;
};
struct ZCMat {
  const zc* p;
  long long ld;
  long long stride;
};
