// Scalar types of the solver: double and float (real kernels) and interleaved complex
// cplx<double> / cplx<float> (Helmholtz).  Complex values are stored as (re, im) pairs,
// the layout of std::complex, MATLAB's interleaved complex API and numpy complex arrays.
// Complex symmetric operators use plain transposes; conj() appears only inside the
// Householder QR (Q^H).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace hssc {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
  cplx() = default;  // trivial, so cplx arrays may live in LDS
  __host__ __device__ constexpr cplx(R r) : re(r), im(0) {}  // NOLINT: real -> complex
  __host__ __device__ constexpr cplx(R r, R i) : re(r), im(i) {}
  template <typename Q, typename = typename std::enable_if<!std::is_same<Q, R>::value>::type>
  __host__ __device__ explicit constexpr cplx(const cplx<Q>& o) : re(R(o.re)), im(R(o.im)) {}
  __host__ __device__ cplx& operator+=(const cplx& o) {
    re += o.re, im += o.im;
    return *this;
  }
  __host__ __device__ cplx& operator-=(const cplx& o) {
    re -= o.re, im -= o.im;
    return *this;
  }
  __host__ __device__ cplx& operator*=(const cplx& o) {
    const R r = re * o.re - im * o.im;
    im = re * o.im + im * o.re;
    re = r;
    return *this;
  }
};

template <typename R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, const cplx<R>& b) {
  return a += b;
}
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, const cplx<R>& b) {
  return a -= b;
}
template <typename R>
__host__ __device__ inline cplx<R> operator-(const cplx<R>& a) {
  return cplx<R>(-a.re, -a.im);
}
template <typename R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, const cplx<R>& b) {
  return a *= b;
}
// Smith's algorithm (no overflow for |b| near the range limits).
template <typename R>
__host__ __device__ inline cplx<R> operator/(const cplx<R>& a, const cplx<R>& b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const R r = b.im / b.re, d = b.re + b.im * r;
    return cplx<R>((a.re + a.im * r) / d, (a.im - a.re * r) / d);
  }
  const R r = b.re / b.im, d = b.im + b.re * r;
  return cplx<R>((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}
template <typename R>
__host__ __device__ inline bool operator==(const cplx<R>& a, const cplx<R>& b) {
  return a.re == b.re && a.im == b.im;
}
template <typename R>
__host__ __device__ inline bool operator!=(const cplx<R>& a, const cplx<R>& b) {
  return !(a == b);
}

using zdouble = cplx<double>;
using zfloat = cplx<float>;

template <typename T>
struct scalar_traits {
  using real = T;
  static constexpr bool is_complex = false;
};
template <typename R>
struct scalar_traits<cplx<R>> {
  using real = R;
  static constexpr bool is_complex = true;
};
template <typename T>
using real_t = typename scalar_traits<T>::real;
template <typename T>
constexpr bool is_complex_v = scalar_traits<T>::is_complex;

// conj, |re| + |im| (LAPACK's cabs1, the pivot magnitude of izamax), |x|^2, Re x
__host__ __device__ inline double conj_s(double x) { return x; }
__host__ __device__ inline float conj_s(float x) { return x; }
template <typename R>
__host__ __device__ inline cplx<R> conj_s(const cplx<R>& x) {
  return cplx<R>(x.re, -x.im);
}
__host__ __device__ inline double abs1(double x) { return fabs(x); }
__host__ __device__ inline float abs1(float x) { return fabsf(x); }
template <typename R>
__host__ __device__ inline R abs1(const cplx<R>& x) {
  return fabs(x.re) + fabs(x.im);
}
__host__ __device__ inline double abs2(double x) { return x * x; }
__host__ __device__ inline float abs2(float x) { return x * x; }
template <typename R>
__host__ __device__ inline R abs2(const cplx<R>& x) {
  return x.re * x.re + x.im * x.im;
}
__host__ __device__ inline double re_s(double x) { return x; }
__host__ __device__ inline float re_s(float x) { return x; }
template <typename R>
__host__ __device__ inline R re_s(const cplx<R>& x) {
  return x.re;
}
// x times a real scalar (two products for complex x, not a complex product)
__host__ __device__ inline double scale_s(double x, double s) { return x * s; }
__host__ __device__ inline float scale_s(float x, float s) { return x * s; }
template <typename R>
__host__ __device__ inline cplx<R> scale_s(const cplx<R>& x, R s) {
  return cplx<R>(x.re * s, x.im * s);
}
__host__ __device__ inline double im_s(double) { return 0.0; }
__host__ __device__ inline float im_s(float) { return 0.0f; }
template <typename R>
__host__ __device__ inline R im_s(const cplx<R>& x) {
  return x.im;
}

// The FP64 type of the same field: double for real T, complex double for complex T.
template <typename T>
using wide_t = typename std::conditional<is_complex_v<T>, cplx<double>, double>::type;

// Wave shuffles for every scalar type.
__device__ inline double shfl_s(double v, int src) { return __shfl(v, src); }
__device__ inline float shfl_s(float v, int src) { return __shfl(v, src); }
template <typename R>
__device__ inline cplx<R> shfl_s(const cplx<R>& v, int src) {
  return cplx<R>(__shfl(v.re, src), __shfl(v.im, src));
}
__device__ inline double shfl_xor_s(double v, int m) { return __shfl_xor(v, m); }
__device__ inline float shfl_xor_s(float v, int m) { return __shfl_xor(v, m); }
template <typename R>
__device__ inline cplx<R> shfl_xor_s(const cplx<R>& v, int m) {
  return cplx<R>(__shfl_xor(v.re, m), __shfl_xor(v.im, m));
}

// Broadcast of lane `l` (wave-uniform) through the scalar unit: v_readlane, no LDS traffic.
__device__ inline float readlane_s(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double readlane_s(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b), l);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <typename R>
__device__ inline cplx<R> readlane_s(const cplx<R>& v, int l) {
  return cplx<R>(readlane_s(v.re, l), readlane_s(v.im, l));
}

// Sum over each aligned group of eight lanes, the result in all eight (DPP moves on the VALU:
// pairs, quads, then the two quads of the group through the half-row mirror).  Every lane
// of a group forms the same sums in the same order up to commutation, so all eight hold the
// same bits.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <int CTRL>
__device__ inline float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL, typename R>
__device__ inline cplx<R> dpp_move(const cplx<R>& v) {
  return cplx<R>(dpp_move<CTRL>(v.re), dpp_move<CTRL>(v.im));
}
template <typename T>
__device__ inline T sum8_lanes(T v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_move<0x141>(v);  // row_half_mirror: lane 7 - lane within the group of eight
  return v;
}

}  // namespace hssc
