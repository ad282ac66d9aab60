// Shared device-side definitions for the MPIDForce gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpid {

// Physical constants of the reference force
// (platforms/reference/src/SimTKReference/MPIDReferenceForce.cpp:39, :49, :2545).
constexpr double kElectric = 138.935455846;
constexpr double kDebye = 48.033324;
constexpr double kSqrtPi = 1.77245385091;
constexpr int kOrder = 6;            // B-spline order, hard-wired in the reference
constexpr int kTile = 64;            // atoms per block = one wavefront
constexpr int kMaxPartners = 16;     // covalent partners with a non-unit scale per atom
constexpr double kFixed = 4294967296.0;  // 2^32: fixed-point scale of every accumulator

typedef unsigned long long acc_t;

// Fixed-point accumulation (deterministic, order independent).
__device__ __forceinline__ void acc_add(acc_t* p, float v) {
    atomicAdd(p, (acc_t)(long long)(v * 4294967296.0f));
}
__device__ __forceinline__ void acc_add(acc_t* p, double v) {
    atomicAdd(p, (acc_t)(long long)(v * kFixed));
}
__host__ __device__ __forceinline__ double acc_get(acc_t v) { return (double)(long long)v / kFixed; }

template <class R> struct V3 {
    R x, y, z;
};
template <class R> __device__ __forceinline__ V3<R> mk(R a, R b, R c) { return V3<R>{a, b, c}; }
template <class R> __device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> __device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> __device__ __forceinline__ V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> __device__ __forceinline__ R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> __device__ __forceinline__ V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ float rsqrt_r(float x) { return rsqrtf(x); }
__device__ __forceinline__ double rsqrt_r(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float erfc_r(float x) { return erfcf(x); }
__device__ __forceinline__ double erfc_r(double x) { return erfc(x); }
__device__ __forceinline__ float exp_r(float x) { return expf(x); }
__device__ __forceinline__ double exp_r(double x) { return exp(x); }

// Rectangular periodic box as seen by the kernels.
template <class R> struct Box {
    R L[3];
    R inv[3];
};

template <class R> __device__ __forceinline__ V3<R> min_image(V3<R> d, const Box<R>& b) {
    d.x -= b.L[0] * floor(d.x * b.inv[0] + R(0.5));
    d.y -= b.L[1] * floor(d.y * b.inv[1] + R(0.5));
    d.z -= b.L[2] * floor(d.z * b.inv[2] + R(0.5));
    return d;
}

// Per-atom multipole record, 20 reals: q, dipole xyz, quadrupole
// (XX XY XZ YY YZ ZZ), octopole (XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ).
constexpr int kMP = 20;

// Scale class of a covalent partner (packed into the low 2 bits of a partner entry).
enum ScaleClass { kScaleZero = 0, kScale14 = 1 };

}  // namespace mpid
