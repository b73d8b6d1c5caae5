// mbr_core.hpp — vector type, precision policy and integer helpers shared by
// the host API and the gfx950 kernels of the render core.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MBR_HD __host__ __device__ __forceinline__

namespace mbr {

template <class T>
struct V3 {
  T x, y, z;
};

template <class T>
MBR_HD V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T>
MBR_HD V3<T> operator+(V3<T> a, V3<T> b) { return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
MBR_HD V3<T> operator-(V3<T> a, V3<T> b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
MBR_HD V3<T> operator*(V3<T> a, T s) { return V3<T>{a.x * s, a.y * s, a.z * s}; }
template <class T>
MBR_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class To, class From>
MBR_HD V3<To> cvt(V3<From> a) { return V3<To>{(To)a.x, (To)a.y, (To)a.z}; }

// Transcendentals per precision.  FP64 uses the device library's double
// routines; FP32 uses its single-precision routines.
template <class T>
struct Math;

template <>
struct Math<double> {
  static MBR_HD double sqrt(double x) { return ::sqrt(x); }
  static MBR_HD double asin(double x) { return ::asin(x); }
  static MBR_HD double acos(double x) { return ::acos(x); }
  static MBR_HD double exp(double x) { return ::exp(x); }
  static MBR_HD double atan2(double y, double x) { return ::atan2(y, x); }
  static MBR_HD double pow(double x, double y) { return ::pow(x, y); }
  static MBR_HD double cos(double x) { return ::cos(x); }
  static MBR_HD double sin(double x) { return ::sin(x); }
  static MBR_HD double log(double x) { return ::log(x); }
  static MBR_HD double fabs(double x) { return ::fabs(x); }
  static MBR_HD double copysign(double a, double b) { return ::copysign(a, b); }
};

template <>
struct Math<float> {
  static MBR_HD float sqrt(float x) { return ::sqrtf(x); }
  static MBR_HD float asin(float x) { return ::asinf(x); }
  static MBR_HD float acos(float x) { return ::acosf(x); }
  static MBR_HD float exp(float x) { return ::expf(x); }
  static MBR_HD float atan2(float y, float x) { return ::atan2f(y, x); }
  static MBR_HD float pow(float x, float y) { return ::powf(x, y); }
  static MBR_HD float cos(float x) { return ::cosf(x); }
  static MBR_HD float sin(float x) { return ::sinf(x); }
  static MBR_HD float log(float x) { return ::logf(x); }
  static MBR_HD float fabs(float x) { return ::fabsf(x); }
  static MBR_HD float copysign(float a, float b) { return ::copysignf(a, b); }
};

template <class T>
MBR_HD T length(V3<T> a) { return Math<T>::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// |a| with every product and sum rounded on its own (no fused multiply-add),
// as the reference's CVector3::Length compiles on x86-64
template <class T>
MBR_HD T length_nc(V3<T> a) {
#pragma clang fp contract(off)
  return Math<T>::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}

// Hardware reciprocal / reciprocal square root / log2 of the FP32 path
// (v_rcp_f32, v_rsq_f32, v_log_f32: about 1 ulp), in place of the
// correctly rounded division and square-root sequences.
// The host versions (used only by host-side checks of these templates) are
// the plain library operations.
MBR_HD float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
MBR_HD float fast_rsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / ::sqrtf(x);
#endif
}
// FP64 forms of the same operations for templates shared by both
// precisions: the correctly rounded division, square root and log.
MBR_HD double fast_rcp(double x) { return 1.0 / x; }
// FP64 1/sqrt: the hardware v_rsq_f64 estimate refined by two Newton steps
// (y <- y (3/2 - x/2 y^2)), about 1 ulp; zero and infinity keep the
// estimate's inf / 0.
MBR_HD double fast_rsq(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y0 = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double y = y0 * (1.5 - h * y0 * y0);
  y = y * (1.5 - h * y * y);
  return (x > 0.0 && x < 1.7976931348623157e308) ? y : y0;
#else
  return 1.0 / Math<double>::sqrt(x);
#endif
}
MBR_HD double fast_ln(double x) { return Math<double>::log(x); }
// a / b: exact in FP64, hardware reciprocal in FP32
MBR_HD double fdiv(double a, double b) { return a / b; }
MBR_HD float fdiv(float a, float b) { return a * fast_rcp(b); }
MBR_HD float fast_ln(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_logf(x) * 0.69314718055994531f;
#else
  return ::logf(x);
#endif
}

// CVector3::Normalize (src/algebra.hpp:104-111): multiply by 1/length
template <class T>
MBR_HD V3<T> normalize(V3<T> a) {
  T n = T(1) / length(a);
  return V3<T>{a.x * n, a.y * n, a.z * n};
}

// Row-major 3x3 product m*v.
template <class T>
MBR_HD V3<T> mat_apply(const T *m, V3<T> v) {
  return V3<T>{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
               m[6] * v.x + m[7] * v.y + m[8] * v.z};
}

// "lowbias32" integer finaliser (C. Wellons, hash-prospector).
MBR_HD uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// Counter-based stand-in for Random(1000) (src/common_math.cpp:25-32) at
// one march step: an integer in [0, 1000].
MBR_HD uint32_t jitter_draw(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t step) {
  uint32_t h = mix32(seed ^ 0x9e3779b9U);
  h = mix32(h ^ frame);
  h = mix32(h ^ pixel);
  h = mix32(h ^ step);
  return h % 1001U;
}

// double -> int as the reference's x86-64 build converts it: values out of
// the int range and NaN become INT_MIN (cvttsd2si), where the GPU's
// conversion would saturate.
MBR_HD int x86_d2i(double d) {
  if (!(d > -2147483649.0 && d < 2147483648.0)) return (int)0x80000000u;
  return (int)d;
}
MBR_HD int x86_abs(int v) { return v == (int)0x80000000u ? v : (v < 0 ? -v : v); }
MBR_HD uint16_t x86_d2u16(double d) { return (uint16_t)(uint32_t)x86_d2i(d); }
MBR_HD uint8_t x86_d2u8(double d) { return (uint8_t)(uint32_t)x86_d2i(d); }

}  // namespace mbr
