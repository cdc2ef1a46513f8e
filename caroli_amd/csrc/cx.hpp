// Complex helpers of the device code: double2 / float2 as complex f64 / f32.
#pragma once
#include <hip/hip_runtime.h>

namespace caroli::gpu {

template <class T>
struct Real;
template <>
struct Real<double2> {
    using type = double;
};
template <>
struct Real<float2> {
    using type = float;
};

template <class T>
__device__ __forceinline__ T cx(double re, double im) {
    using R = typename Real<T>::type;
    T t;
    t.x = R(re);
    t.y = R(im);
    return t;
}
template <class T>
__device__ __forceinline__ T cmul(T a, T b) {
    T c;
    c.x = a.x * b.x - a.y * b.y;
    c.y = a.x * b.y + a.y * b.x;
    return c;
}
template <class T>
__device__ __forceinline__ T cadd(T a, T b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
template <class T>
__device__ __forceinline__ T csub(T a, T b) {
    a.x -= b.x;
    a.y -= b.y;
    return a;
}
template <class T>
__device__ __forceinline__ T cconj(T a) {
    a.y = -a.y;
    return a;
}
template <class T>
__device__ __forceinline__ double cabs(T a) {
    return hypot(double(a.x), double(a.y));
}
// Squared modulus: the pivot searches compare these (the order of the moduli, no square root).
template <class T>
__device__ __forceinline__ double cabs2(T a) {
    return double(a.x) * double(a.x) + double(a.y) * double(a.y);
}
template <class T>
__device__ __forceinline__ T crecip(T a) {
    // 1 / a with one division of the squared modulus, scaled against overflow.
    using R = typename Real<T>::type;
    R s = fmax(fabs(a.x), fabs(a.y));
    R ar = a.x / s, ai = a.y / s;
    R d = (ar * ar + ai * ai) * s;
    T r;
    r.x = ar / d;
    r.y = -ai / d;
    return r;
}
template <class TO, class TI>
__device__ __forceinline__ TO convert(TI a) {
    using R = typename Real<TO>::type;
    TO t;
    t.x = R(a.x);
    t.y = R(a.y);
    return t;
}

}  // namespace caroli::gpu
