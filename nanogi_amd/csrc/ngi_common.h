// Shared host/device math and random-number helpers.
//
// Vector math follows the glm usage of include/nanogi/rt.hpp (dvec3 in the
// reference; here a template so the same code runs in FP32 and FP64).
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define NGI_HD __host__ __device__ __forceinline__
#define NGI_UNROLL _Pragma("unroll")
#else
#define NGI_HD inline
#define NGI_UNROLL
#endif

namespace ngi {

template <typename T>
struct V3 {
    T x, y, z;
    NGI_HD V3() : x(0), y(0), z(0) {}
    NGI_HD V3(T a, T b, T c) : x(a), y(b), z(c) {}
    NGI_HD explicit V3(T a) : x(a), y(a), z(a) {}
    template <typename U>
    NGI_HD explicit V3(const V3<U>& o) : x((T)o.x), y((T)o.y), z((T)o.z) {}
    NGI_HD T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    NGI_HD T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

template <typename T> NGI_HD V3<T> operator+(V3<T> a, V3<T> b) { return V3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> NGI_HD V3<T> operator-(V3<T> a, V3<T> b) { return V3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> NGI_HD V3<T> operator-(V3<T> a) { return V3<T>(-a.x, -a.y, -a.z); }
template <typename T> NGI_HD V3<T> operator*(V3<T> a, V3<T> b) { return V3<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <typename T> NGI_HD V3<T> operator*(V3<T> a, T s) { return V3<T>(a.x * s, a.y * s, a.z * s); }
template <typename T> NGI_HD V3<T> operator*(T s, V3<T> a) { return V3<T>(a.x * s, a.y * s, a.z * s); }
template <typename T> NGI_HD V3<T> operator/(V3<T> a, T s) { return V3<T>(a.x / s, a.y / s, a.z / s); }
template <typename T> NGI_HD V3<T> operator/(V3<T> a, V3<T> b) { return V3<T>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <typename T> NGI_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> NGI_HD V3<T> cross(V3<T> a, V3<T> b) {
    return V3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> NGI_HD T length(V3<T> a) { return sqrt(dot(a, a)); }
template <typename T> NGI_HD V3<T> normalize(V3<T> a) { return a / length(a); }
template <typename T> NGI_HD bool is_zero(V3<T> a) { return a.x == T(0) && a.y == T(0) && a.z == T(0); }

typedef V3<float> f3;
typedef V3<double> d3;

// ---------------------------------------------------------------------------
// Philox-4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Round function:
// two 32x32->64 multiplies per round, Weyl-sequence key schedule.
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

NGI_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    NGI_UNROLL
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// 24-bit uniform in [0, 1): exact in float and in double, so FP32 and FP64
// paths consume bit-identical random numbers.
NGI_HD float u24f(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
NGI_HD double u24d(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// Random-stream layout shared by every engine and by the CPU oracle.
//   key     = (seed, stream_id)
//   counter = (path_lo, path_hi, vertex, stream)
// In per-pixel sampling path_lo = pixel index and path_hi = sample index of
// that pixel; in reference-uniform sampling (path_lo, path_hi) is the 64-bit
// global sample index with bit 31 of path_hi set.
// vertex = numVertices at the point of use (1 = camera vertex).
// stream 0 = extension: (u.x, u.y, uComp, uRR)  (src/nanogi.cpp:495, 584)
// stream 1 = next-event estimation: (uLight, u2.x, u2.y, -)  (src/nanogi.cpp:659, 670)
enum { NGI_STREAM_EXT = 0, NGI_STREAM_NEE = 1 };
#define NGI_UNIFORM_FLAG 0x80000000u

}  // namespace ngi
