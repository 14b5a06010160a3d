// Shared device helpers for the MNPBEM HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mnp {

// interleaved complex number (re, im)
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) { cx<T> z; z.re = r; z.im = i; return z; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return mk(s * a.re, s * a.im); }
// a - b*c
template <typename T>
__device__ __forceinline__ cx<T> fms(cx<T> a, cx<T> b, cx<T> c) {
    return mk(a.re - (b.re * c.re - b.im * c.im), a.im - (b.re * c.im + b.im * c.re));
}
template <typename T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
    // Smith's algorithm keeps the scaling well behaved for wide-range values
    if (fabs(b.re) >= fabs(b.im)) {
        T r = b.im / b.re, d = b.re + b.im * r;
        return mk((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    } else {
        T r = b.re / b.im, d = b.re * r + b.im;
        return mk((a.re * r + a.im) / d, (a.im * r - a.re) / d);
    }
}
template <typename T>
__device__ __forceinline__ cx<T> crecip(cx<T> b) { return cdiv(mk<T>(1, 0), b); }
template <typename T>
__device__ __forceinline__ T cabs1(cx<T> a) { return fabs(a.re) + fabs(a.im); }

template <typename T>
__device__ __forceinline__ cx<T> conv(cx<double> a) { return mk<T>((T)a.re, (T)a.im); }

// exp(i k d) for complex k: e^{-Im(k) d} (cos(Re(k) d) + i sin(Re(k) d))
__device__ __forceinline__ cx<double> expikd(cx<double> k, double d) {
    double s, c;
    sincos(k.re * d, &s, &c);
    double m = exp(-k.im * d);
    return mk(m * c, m * s);
}

}  // namespace mnp
