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
// B-spline order of the PME spreading and gathering: 6, as the reference
// hard-wires (MPIDReferenceForce.cpp:2539; MPIDForce::getPmeBSplineOrder), or 5
// in a build with -DMPID_PME_ORDER=5 (libmpid_hip_order5.so).  The tile
// geometry of the spreading (window origins 6 points before a tile, windows
// of T + 8 points, owned cores of T >= 8) covers both orders.
#ifndef MPID_PME_ORDER
#define MPID_PME_ORDER 6
#endif
constexpr int kOrder = MPID_PME_ORDER;
static_assert(kOrder == 5 || kOrder == 6, "B-spline order 5 or 6");
constexpr int kTile = 64;            // atoms per block = one wavefront
constexpr double kFixed = 4294967296.0;  // 2^32: fixed-point scale of every accumulator

typedef unsigned long long acc_t;

// Generic (templated) arithmetic helpers and pair formulas are host+device so
// that they can also be instantiated on the host with an operation-counting
// number type; the float/double overloads of the math functions are device-only.
#define MPID_HD __host__ __device__ __forceinline__

// Fixed-point accumulation (deterministic, order independent).
__device__ __forceinline__ void acc_add(acc_t* p, float v) {
    atomicAdd(p, (acc_t)(long long)(v * 4294967296.0f));
}
__device__ __forceinline__ void acc_add(acc_t* p, double v) {
    atomicAdd(p, (acc_t)(long long)(v * kFixed));
}
__host__ __device__ __forceinline__ double acc_get(acc_t v) { return (double)(long long)v / kFixed; }

// Workgroups are dispatched to the chip's 8 XCDs in turn (workgroup b to
// XCD b mod 8), and each XCD has its own L2.  Renumbering b so that each XCD
// takes one contiguous eighth of a grid keeps neighbouring work items (sorted
// atoms and their stencils, block rows) on one L2.  A bijection of [0, nb).
// Used where it measured faster: the gathers and the induced-field kernel
// (the tile, mask, list-fill and owner-spread kernels ran no faster with it).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// Kernels of a solver iteration take the solver's stop flag: the iteration
// was issued ahead of the host's convergence test and the solve has stopped,
// so the kernel returns at once (before any barrier; the flag is uniform).
__device__ __forceinline__ bool skipped(const double* flag) { return flag != nullptr && *flag != 0.0; }

template <class R> struct V3 {
    R x, y, z;
};
template <class R> MPID_HD V3<R> mk(R a, R b, R c) { return V3<R>{a, b, c}; }
template <class R> MPID_HD V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> MPID_HD V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> MPID_HD V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> MPID_HD R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> MPID_HD V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Precision-dispatched math.  FP32 uses the hardware reciprocal and
// reciprocal square root (1 ulp; pair distances are never denormal) and a
// fitted erfc; FP64 uses the library functions (also on the host, where the
// pair formulas are instantiated for operation counts and identity checks).
__device__ __forceinline__ float rsqrt_r(float x) { return __builtin_amdgcn_rsqf(x); }
MPID_HD double rsqrt_r(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float div_r(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
MPID_HD double div_r(double a, double b) { return a / b; }
__device__ __forceinline__ float exp_r(float x) { return expf(x); }
MPID_HD double exp_r(double x) { return exp(x); }

// erfc(x) for x >= 0, given e = exp(-x^2) (which the Ewald terms need anyway).
// FP32: erfc(x) = e * t * Q(t), t = 1 / (1 + 0.485 x), with Q a degree-8
// polynomial fitted (iteratively reweighted least squares, minimax-like) to
// exp(x^2) erfc(x) / t on 0 <= x <= 9: relative error 1.7e-8 in exact
// arithmetic, below 4e-7 when evaluated in FP32.
constexpr float kErfcP = 0.485f;
constexpr float kErfcQ[9] = {0.2735917531303337f,  0.27457141080238046f, 0.23205431683139313f,
                             0.22971345991274555f, -0.08908535117725187f, 0.39934447387421285f,
                             -0.5558771579615225f, 0.29188334026310675f, -0.05619626121834402f};
__device__ __forceinline__ float erfc_r(float x, float e) {
    const float t = __builtin_amdgcn_rcpf(1.0f + kErfcP * x);
    float q = kErfcQ[8];
#pragma unroll
    for (int k = 7; k >= 0; k--) q = fmaf(q, t, kErfcQ[k]);
    return e * t * q;
}
MPID_HD double erfc_r(double x, double) { return erfc(x); }

// Periodic box as seen by the kernels, in OpenMM's reduced triclinic form:
// box vectors a = (L0, 0, 0), b = (bx, L1, 0), c = (cx, cy, L2) with
// |bx|, |cx| <= L0/2 and |cy| <= L1/2 (bx = cx = cy = 0: rectangular).
template <class R> struct Box {
    R L[3];
    R inv[3];  // 1 / L_k
    R bx, cx, cy;
};

// Nearest image of a separation: the c, b and a shifts in turn, as
// getPeriodicDelta does (MPIDReferenceForce.cpp:2671-2676).  For a
// rectangular box the off-diagonal terms vanish and each axis wraps alone.
template <class R> MPID_HD V3<R> min_image(V3<R> d, const Box<R>& b) {
    const R nc = floor(d.z * b.inv[2] + R(0.5));
    d.z -= nc * b.L[2];
    d.y -= nc * b.cy;
    d.x -= nc * b.cx;
    const R nb = floor(d.y * b.inv[1] + R(0.5));
    d.y -= nb * b.L[1];
    d.x -= nb * b.bx;
    d.x -= b.L[0] * floor(d.x * b.inv[0] + R(0.5));
    return d;
}

// Fractional coordinates (along a, b, c) of a position: r = s0 a + s1 b + s2 c.
template <class R> MPID_HD V3<R> to_frac(V3<R> r, const Box<R>& b) {
    const R s2 = r.z * b.inv[2];
    const R s1 = (r.y - s2 * b.cy) * b.inv[1];
    const R s0 = (r.x - s1 * b.bx - s2 * b.cx) * b.inv[0];
    return V3<R>{s0, s1, s2};
}
template <class R> MPID_HD V3<R> from_frac(V3<R> s, const Box<R>& b) {
    return V3<R>{s.x * b.L[0] + s.y * b.bx + s.z * b.cx, s.y * b.L[1] + s.z * b.cy, s.z * b.L[2]};
}

// Per-atom multipole record, 20 reals: q, dipole xyz, quadrupole
// (XX XY XZ YY YZ ZZ), octopole (XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ).
constexpr int kMP = 20;

// Scale class of a covalent partner (packed into the low 2 bits of a partner entry).
enum ScaleClass { kScaleZero = 0, kScale14 = 1 };

}  // namespace mpid
