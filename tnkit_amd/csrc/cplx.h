// Complex element type for the complex128 / complex64 forms of the path (real and imaginary parts
// interleaved, the layout of torch.complex128 and numpy.complex128).
#pragma once
#include <hip/hip_runtime.h>

namespace tnh {

template <typename T>
struct cpx {
  T re, im;
  __host__ __device__ cpx() : re(0), im(0) {}
  __host__ __device__ cpx(T r) : re(r), im(0) {}  // NOLINT: implicit from a real
  __host__ __device__ cpx(T r, T i) : re(r), im(i) {}
  template <typename U>
  __host__ __device__ explicit cpx(const cpx<U>& o) : re(static_cast<T>(o.re)), im(static_cast<T>(o.im)) {}
};

template <typename T>
__host__ __device__ inline cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ inline cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ inline cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cpx<T>& operator+=(cpx<T>& a, cpx<T> b) {
  a.re += b.re;
  a.im += b.im;
  return a;
}
template <typename T>
__host__ __device__ inline bool operator!=(cpx<T> a, cpx<T> b) { return a.re != b.re || a.im != b.im; }
template <typename T>
__host__ __device__ inline cpx<T> conj(cpx<T> a) { return {a.re, -a.im}; }

typedef cpx<double> c128;
typedef cpx<float> c64;

}  // namespace tnh
