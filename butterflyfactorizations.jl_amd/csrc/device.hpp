// device.hpp -- complex arithmetic and wavefront reductions shared by the
// device code (kernels.hip, recompress.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace bfg {

template <typename R> struct C { R x, y; };

template <typename R> __device__ __forceinline__ C<R> cmk(R a, R b) { return C<R>{a, b}; }
template <typename R> __device__ __forceinline__ C<R> cadd(C<R> a, C<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> __device__ __forceinline__ C<R> csub(C<R> a, C<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> __device__ __forceinline__ C<R> cmul(C<R> a, C<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <typename R> __device__ __forceinline__ C<R> cmulc(C<R> a, C<R> b) {
    return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
// acc += a * b
template <typename R> __device__ __forceinline__ void cfma(C<R> &acc, C<R> a, C<R> b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a) * b
template <typename R> __device__ __forceinline__ void cfmac(C<R> &acc, C<R> a, C<R> b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace bfg
