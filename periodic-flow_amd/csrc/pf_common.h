// Shared device helpers: complex arithmetic, launch geometry, constants.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf {

template <typename T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };
template <typename T> using cplx = typename C2<T>::type;

template <typename T> __host__ __device__ __forceinline__ cplx<T> mk(T r, T i) {
  cplx<T> z; z.x = r; z.y = i; return z;
}
template <typename T> __device__ __forceinline__ cplx<T> add(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cplx<T> sub(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cplx<T> mul(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ cplx<T> scl(cplx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T> __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return mk<T>(a.x, -a.y); }
// multiply by i
template <typename T> __device__ __forceinline__ cplx<T> mul_i(cplx<T> a) { return mk<T>(-a.y, a.x); }
// multiply by S*i with S = +1 or -1
template <int S, typename T> __device__ __forceinline__ cplx<T> mul_si(cplx<T> a) {
  return S > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}
template <typename T> __device__ __forceinline__ T abs2(cplx<T> a) { return a.x * a.x + a.y * a.y; }

// Global loads and stores of the passes' streams.  Each element of a pass's
// input and output is touched once per pass, so a launch variant may give
// them the nontemporal hint (the "nt" cache policy of global_load /
// global_store): streamed lines then bypass the write-back caching that a
// working set larger than the last-level cache gains nothing from.  NT: the
// hint on (a compile-time choice of the launch variant).
template <bool NT, typename T> __device__ __forceinline__ void st_stream(cplx<T>* p, cplx<T> v) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (v2*)p);
  } else {
    *p = v;
  }
}
template <bool NT, typename T> __device__ __forceinline__ cplx<T> ld_stream(const cplx<T>* p) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 w = __builtin_nontemporal_load((const v2*)p);
    return mk<T>(w.x, w.y);
  } else {
    return *p;
  }
}
// Cache policy of a launch: bit 0 nontemporal stores, bit 1 nontemporal loads.
constexpr bool cp_store(int cp) { return (cp & 1) != 0; }
constexpr bool cp_load(int cp) { return (cp & 2) != 0; }

// Signed integer mode number of index i on a complex axis of length n
// (numpy fftfreq convention: the Nyquist index n/2 maps to -n/2).
__host__ __device__ __forceinline__ int cmode(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

}  // namespace pf
