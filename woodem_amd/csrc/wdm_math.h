// Small fixed-size algebra for the timestep kernels.  Every 3-term sum is
// evaluated left to right, (a0*b0 + a1*b1) + a2*b2, and the library is compiled
// with -ffp-contract=off, so a value computed here rounds exactly like the same
// expression evaluated on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define WDM_HD __host__ __device__ __forceinline__

template <typename T>
struct V3T {
    T x, y, z;
};
typedef V3T<double> D3;
typedef V3T<float> F3;

template <typename T> WDM_HD V3T<T> mk3(T a, T b, T c) { V3T<T> r; r.x = a; r.y = b; r.z = c; return r; }
template <typename T> WDM_HD V3T<T> operator+(const V3T<T>& a, const V3T<T>& b) { return mk3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> WDM_HD V3T<T> operator-(const V3T<T>& a, const V3T<T>& b) { return mk3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> WDM_HD V3T<T> operator*(const V3T<T>& a, T s) { return mk3<T>(a.x * s, a.y * s, a.z * s); }
template <typename T> WDM_HD V3T<T> operator/(const V3T<T>& a, T s) { return mk3<T>(a.x / s, a.y / s, a.z / s); }
template <typename T> WDM_HD T dot3(const V3T<T>& a, const V3T<T>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <typename T> WDM_HD T sqn3(const V3T<T>& a) { return dot3(a, a); }
template <typename T> WDM_HD V3T<T> cross3(const V3T<T>& a, const V3T<T>& b) {
    return mk3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> WDM_HD bool iszero3(const V3T<T>& a) { return a.x == 0 && a.y == 0 && a.z == 0; }
template <typename T> WDM_HD T comp3(const V3T<T>& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <typename T> WDM_HD void setcomp3(V3T<T>& a, int i, T v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }
template <typename T> WDM_HD V3T<T> abs3(const V3T<T>& a) { return mk3<T>(fabs(a.x), fabs(a.y), fabs(a.z)); }

WDM_HD double wsqrt(double x) { return sqrt(x); }
WDM_HD float wsqrt(float x) { return sqrtf(x); }

// v / |v| when |v|^2 > 0 (division by the root, not multiplication by its inverse)
template <typename T> WDM_HD V3T<T> normalized3(const V3T<T>& a) {
    T z = sqn3(a);
    if (z > T(0)) return a / wsqrt(z);
    return a;
}

template <typename T> WDM_HD V3T<T> cvt3(const D3& a) { return mk3<T>((T)a.x, (T)a.y, (T)a.z); }
WDM_HD D3 todbl(const F3& a) { return mk3<double>((double)a.x, (double)a.y, (double)a.z); }
WDM_HD D3 todbl(const D3& a) { return a; }

WDM_HD D3 xyz(const double4& v) { return mk3<double>(v.x, v.y, v.z); }

// 3x3 matrix, row-major, used for the periodic cell
struct M3 {
    double m[9];
};
WDM_HD D3 mv3(const M3& a, const D3& v) {
    return mk3<double>((a.m[0] * v.x + a.m[1] * v.y) + a.m[2] * v.z,
                       (a.m[3] * v.x + a.m[4] * v.y) + a.m[5] * v.z,
                       (a.m[6] * v.x + a.m[7] * v.y) + a.m[8] * v.z);
}

// sign function of the reference's Mathr::Sign: -1, 0, +1
WDM_HD double sgn(double f) { return f < 0 ? -1. : (f > 0 ? 1. : 0.); }
