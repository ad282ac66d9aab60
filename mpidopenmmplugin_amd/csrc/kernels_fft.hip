// 3D FFT and influence-function convolution of the PME reciprocal sum.
//
// The charge grid Q is real, [n0][n1][n2] with z fastest (the reference's
// x*ny*nz + y*nz + z order, MPIDReferenceForce.cpp:3269-3327).  Its spectrum
// is kept as the half kz < n2/2+1: the grid is real and the influence
// function even, so the other half is the complex conjugate and is never
// stored.  The half is stored plane by plane, S[kz][x][y], so that every pass
// moves contiguous runs.  One reciprocal cycle is five passes over the grid:
//
//   1. z, real to half-complex   Q -> S      (forward)
//   2. y, complex, in place      S           (forward)
//   3. x, in place: forward transform, multiply by the influence function
//      (performMPIDReciprocalConvolution, MPIDReferenceForce.cpp:3329-3366),
//      backward transform, all while the lines sit in LDS
//   4. y, complex, in place      S           (backward)
//   5. z, half-complex to real   S -> Q      (backward)
//
// Like the reference's fftpack transforms, neither direction is normalised;
// the convolution constants carry the scale.
//
// Every pass loads a batch of T lines into LDS, runs Stockham stages of radix
// 4, 2, 3 and 5 (twiddles staged in LDS) and writes the lines back.  Index
// arithmetic uses precomputed multiply-shift division by the stage constants.
#include <algorithm>

#include "engine.h"

namespace mpid {

template <class R> struct Cx {
    R x, y;
};
template <class R> __device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class R> __device__ __forceinline__ Cx<R> cadd(Cx<R> a, Cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> __device__ __forceinline__ Cx<R> csub(Cx<R> a, Cx<R> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by s*i, s = +-1
template <class R> __device__ __forceinline__ Cx<R> cmuli(Cx<R> a, R s) { return {-s * a.y, s * a.x}; }

template <class R, int RAD> __device__ __forceinline__ void dft_small(Cx<R>* v, R s) {
    if (RAD == 2) {
        const Cx<R> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (RAD == 4) {
        const Cx<R> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        const Cx<R> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmuli(csub(a1, a3), s);
        v[0] = cadd(t0, t2);
        v[1] = cadd(t1, t3);
        v[2] = csub(t0, t2);
        v[3] = csub(t1, t3);
    } else if (RAD == 3) {
        const R h = R(0.86602540378443865);
        const Cx<R> a0 = v[0], t1 = cadd(v[1], v[2]), t2 = cmuli(csub(v[1], v[2]), s * h);
        const Cx<R> m = {a0.x - R(0.5) * t1.x, a0.y - R(0.5) * t1.y};
        v[0] = cadd(a0, t1);
        v[1] = cadd(m, t2);
        v[2] = csub(m, t2);
    } else if (RAD == 5) {
        const R c1 = R(0.30901699437494742), c2 = R(-0.80901699437494742);
        const R s1 = R(0.95105651629515357), s2 = R(0.58778525229247313);
        const Cx<R> a0 = v[0];
        const Cx<R> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        const Cx<R> m1 = {a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y};
        const Cx<R> m2 = {a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y};
        const Cx<R> n1 = cmuli(Cx<R>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, s);
        const Cx<R> n2 = cmuli(Cx<R>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, s);
        v[0] = cadd(a0, cadd(t1, t2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
}

// One Stockham stage on T lines of length n: butterfly j of line t reads
// in[t][j + r nb] (r < RAD), twiddles by w^(kk r), and writes
// out[t][(j / Ns) Ns RAD + kk + r Ns], kk = j mod Ns.
template <class R, int RAD>
__device__ __forceinline__ void stockham_stage(const Cx<R>* in, Cx<R>* out, int n, int T, const FftStage& st,
                                               const Cx<R>* tw, R s) {
    const int nb = (int)st.nb.d, Ns = (int)st.ns.d;
    for (int b = threadIdx.x; b < T * nb; b += blockDim.x) {
        const int t = (int)fdiv((unsigned)b, st.nb), j = b - t * nb;
        const int jq = (int)fdiv((unsigned)j, st.ns), kk = j - jq * Ns;
        const Cx<R>* li = in + t * n;
        Cx<R>* lo = out + t * n;
        Cx<R> v[RAD];
#pragma unroll
        for (int r = 0; r < RAD; r++) {
            Cx<R> a = li[j + r * nb];
            if (r > 0 && kk > 0) {
                // table: exp(-2 pi i k / n); exp(s 2 pi i k / n) = (w.x, -s w.y)
                const Cx<R> w = tw[kk * r * st.twstep];
                a = cmul(a, Cx<R>{w.x, -s * w.y});
            }
            v[r] = a;
        }
        dft_small<R, RAD>(v, s);
        const int o = jq * Ns * RAD + kk;
#pragma unroll
        for (int r = 0; r < RAD; r++) lo[o + r * Ns] = v[r];
    }
}

// All stages of one axis; returns the buffer that holds the result.
template <class R>
__device__ __forceinline__ Cx<R>* fft_in_lds(Cx<R>* a, Cx<R>* b, const FftAxis& ax, int T, const Cx<R>* tw, R s) {
    Cx<R>* in = a;
    Cx<R>* out = b;
    for (int k = 0; k < ax.nst; k++) {
        const FftStage& st = ax.st[k];
        switch (st.rad) {
            case 4: stockham_stage<R, 4>(in, out, ax.n, T, st, tw, s); break;
            case 2: stockham_stage<R, 2>(in, out, ax.n, T, st, tw, s); break;
            case 3: stockham_stage<R, 3>(in, out, ax.n, T, st, tw, s); break;
            default: stockham_stage<R, 5>(in, out, ax.n, T, st, tw, s); break;
        }
        __syncthreads();
        Cx<R>* t = in;
        in = out;
        out = t;
    }
    return in;
}

template <class R> __device__ __forceinline__ void stage_twiddles(Cx<R>* tw_s, const R2<R>* tw, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) tw_s[k] = Cx<R>{tw[k].x, tw[k].y};
}

// A batch holds T*n <= kFftLdsBytes / (2 * sizeof(complex)) elements (two
// line buffers), at most max_per_thread() per thread.  Loads go to registers
// first, all of them in flight at once, and then to LDS.  T = 16 FP32 or 8
// FP64 lines of 128 make every global access a run of at least 128
// contiguous bytes (see the spectrum layout below).  A pass runs as few
// workgroups as there are batches, so its time is one workgroup's latency:
// the workgroup has one thread per radix-4 butterfly of a stage (512 FP32,
// 256 FP64 threads for n = 128), so that each stage is one step per thread.
constexpr int kFftLdsBytes = 32768;
constexpr int kFftMaxN = 1024;  // largest grid dimension accepted (one line per batch at worst)
template <class R> constexpr int fft_threads() { return sizeof(R) == 4 ? 512 : 256; }
template <class R> constexpr int max_per_thread() {
    constexpr int budget = kFftLdsBytes / (2 * 2 * (int)sizeof(R));
    constexpr int elems = budget > kFftMaxN ? budget : kFftMaxN;
    return (elems + fft_threads<R>() - 1) / fft_threads<R>();
}

// Spectrum layout: S[kz][l], l = x * n1 + y, kz < n2 / 2 + 1 (N01 = n0 * n1
// words per plane).  Pass 1 writes, for each kz, T consecutive lines l as one
// contiguous run; pass 2 (lines along y) reads and writes whole contiguous
// blocks of T lines; pass 3 (lines along x) reads runs of T consecutive y.

// Pass 1 (dir = -1): lines l along z, real Q[l][z] -> S[kz][l].
// Pass 5 (dir = +1): S[kz][l] -> real Q[l][z] (Hermitian completion).
// A batch is T consecutive lines l0 .. l0 + T - 1.
template <class R>
__global__ __launch_bounds__(fft_threads<R>()) void k_fft_z(R* __restrict__ Q, R2<R>* __restrict__ S, FftAxis ax, int T,
                                                       int Tshift, const R2<R>* __restrict__ tw, int nlines, int dir,
                                                       const double* skip) {
    if (skipped(skip)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = ax.n;
    Cx<R>* A = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* B = A + (size_t)T * n;
    Cx<R>* W = B + (size_t)T * n;
    const int n2h = n / 2 + 1;
    const int l0 = blockIdx.x * T;
    const int count = min(T, nlines - l0);
    stage_twiddles(W, tw, n);
    constexpr int MP = max_per_thread<R>();
    const int total = T * n;
    if (dir < 0) {
        const R* src = Q + (size_t)l0 * n;
        const int lim = count * n;
        R v[MP];
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            v[u] = idx < lim ? src[idx] : R(0);
        }
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            if (idx < total) A[idx] = Cx<R>{v[u], R(0)};
        }
    } else {
        // element (t, k) of the batch, t fastest: runs of T along l
        R2<R> v[MP];
        bool cj[MP];
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            const int k = idx >> Tshift, t = idx & (T - 1);
            cj[u] = k >= n2h;
            v[u] = R2<R>{R(0), R(0)};
            if (idx < total && t < count) v[u] = S[(size_t)(cj[u] ? n - k : k) * nlines + l0 + t];
        }
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            const int k = idx >> Tshift, t = idx & (T - 1);
            if (idx < total) A[t * n + k] = Cx<R>{v[u].x, cj[u] ? -v[u].y : v[u].y};
        }
    }
    __syncthreads();
    const Cx<R>* res = fft_in_lds<R>(A, B, ax, T, W, (R)dir);
    if (dir < 0) {
        for (int idx = threadIdx.x; idx < T * n2h; idx += blockDim.x) {
            const int k = idx >> Tshift, t = idx & (T - 1);
            if (t < count) {
                const Cx<R> v = res[t * n + k];
                S[(size_t)k * nlines + l0 + t] = R2<R>{v.x, v.y};
            }
        }
    } else {
        R* dst = Q + (size_t)l0 * n;
        for (int idx = threadIdx.x; idx < T * n; idx += blockDim.x) {
            const int t = (int)fdiv((unsigned)idx, ax.nd);
            if (t < count) dst[idx] = res[idx].x;
        }
    }
}

// Passes 2 and 4: lines along y, a batch = T consecutive x of one kz plane,
// a contiguous block of T * n1 words.
template <class R>
__global__ __launch_bounds__(fft_threads<R>()) void k_fft_y(R2<R>* __restrict__ S, FftAxis ax, int T,
                                                       const R2<R>* __restrict__ tw, int n0, int nxb, int dir,
                                                       const double* skip) {
    if (skipped(skip)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = ax.n;
    Cx<R>* A = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* B = A + (size_t)T * n;
    Cx<R>* W = B + (size_t)T * n;
    const int kz = blockIdx.x / nxb, x0 = (blockIdx.x - kz * nxb) * T;
    const int count = min(T, n0 - x0);
    R2<R>* base = S + ((size_t)kz * n0 + x0) * n;
    stage_twiddles(W, tw, n);
    constexpr int MP = max_per_thread<R>();
    const int lim = count * n;
    {
        R2<R> v[MP];
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            v[u] = idx < lim ? base[idx] : R2<R>{R(0), R(0)};
        }
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            if (idx < T * n) A[idx] = Cx<R>{v[u].x, v[u].y};
        }
    }
    __syncthreads();
    const Cx<R>* res = fft_in_lds<R>(A, B, ax, T, W, (R)dir);
    for (int idx = threadIdx.x; idx < lim; idx += blockDim.x) {
        const Cx<R> v = res[idx];
        base[idx] = R2<R>{v.x, v.y};
    }
}

// Pass 3: lines along x (a batch = T consecutive y of one kz plane, runs of T
// words): forward transform, influence function, backward transform.
template <class R>
__global__ __launch_bounds__(fft_threads<R>()) void k_fft_x_convolve(R2<R>* __restrict__ S, FftAxis ax, int T, int Tshift,
                                                                const R2<R>* __restrict__ tw, int n1, int n2, int nyb,
                                                                ConvParams<R> cp) {
    if (skipped(cp.skip)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = ax.n;
    Cx<R>* A = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* B = A + (size_t)T * n;
    Cx<R>* W = B + (size_t)T * n;
    const int kz = blockIdx.x / nyb, y0 = (blockIdx.x - kz * nyb) * T;
    const int count = min(T, n1 - y0);
    R2<R>* base = S + (size_t)kz * n * n1 + y0;
    stage_twiddles(W, tw, n);
    constexpr int MP = max_per_thread<R>();
    const int total = T * n;
    {
        R2<R> v[MP];
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            const int k = idx >> Tshift, t = idx & (T - 1);
            v[u] = (idx < total && t < count) ? base[(size_t)k * n1 + t] : R2<R>{R(0), R(0)};
        }
#pragma unroll
        for (int u = 0; u < MP; u++) {
            const int idx = threadIdx.x + u * fft_threads<R>();
            const int k = idx >> Tshift, t = idx & (T - 1);
            if (idx < total) A[t * n + k] = Cx<R>{v[u].x, v[u].y};
        }
    }
    __syncthreads();
    Cx<R>* res = fft_in_lds<R>(A, B, ax, T, W, R(-1));
    // the influence function, from its table ([kz][ky][kx])
    const R* gz = cp.infl + (size_t)kz * n1 * n;
    for (int idx = threadIdx.x; idx < T * n; idx += blockDim.x) {
        const int kx = idx >> Tshift, t = idx & (T - 1);
        const int y = y0 + t;
        const R et = t < count ? gz[(size_t)y * n + kx] : R(0);
        Cx<R>& v = res[t * n + kx];
        v.x *= et;
        v.y *= et;
    }
    __syncthreads();
    Cx<R>* other = (res == A) ? B : A;
    const Cx<R>* out = fft_in_lds<R>(res, other, ax, T, W, R(1));
    for (int idx = threadIdx.x; idx < T * n; idx += blockDim.x) {
        const int k = idx >> Tshift, t = idx & (T - 1);
        if (t < count) {
            const Cx<R> v = out[t * n + k];
            base[(size_t)k * n1 + t] = R2<R>{v.x, v.y};
        }
    }
}

// ---------------------------------------------------------------------------
// Small planes (two n0 x n1 buffers and both twiddle tables within 64 KB of
// LDS): passes 2-4 fused, one workgroup per kz plane.  The plane is loaded as
// n0 lines along y, transformed, transposed to n1 lines along x, transformed,
// multiplied by the influence function (as k_fft_x_convolve), transformed
// back along x, transposed back, transformed back along y and stored: one
// dispatch instead of three for the grids of small systems.
template <class R> __host__ __device__ constexpr size_t plane_small_lds(int n0, int n1) {
    return (2 * (size_t)n0 * n1 + n0 + n1) * sizeof(Cx<R>);
}
constexpr size_t kPlaneSmallLds = 65536;

template <class R>
__global__ __launch_bounds__(fft_threads<R>()) void k_fft_plane_small(R2<R>* __restrict__ S, FftAxis ax0, FftAxis ax1,
                                                                 const R2<R>* __restrict__ tw0,
                                                                 const R2<R>* __restrict__ tw1, int n2,
                                                                 ConvParams<R> cp) {
    if (skipped(cp.skip)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n0 = ax0.n, n1 = ax1.n, np = n0 * n1;
    Cx<R>* A = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* B = A + np;
    Cx<R>* W0 = B + np;
    Cx<R>* W1 = W0 + n0;
    const int kz = blockIdx.x;
    R2<R>* base = S + (size_t)kz * np;
    for (int idx = threadIdx.x; idx < np; idx += blockDim.x) {
        const R2<R> v = base[idx];
        A[idx] = Cx<R>{v.x, v.y};
    }
    stage_twiddles(W0, tw0, n0);
    stage_twiddles(W1, tw1, n1);
    __syncthreads();
    // y forward: n0 lines [x][y]
    Cx<R>* res = fft_in_lds<R>(A, B, ax1, n0, W1, R(-1));
    Cx<R>* other = (res == A) ? B : A;
    // transpose [x][y] -> [y][x]
    for (int idx = threadIdx.x; idx < np; idx += blockDim.x) {
        const int x = idx / n1, y = idx - x * n1;
        other[y * n0 + x] = res[idx];
    }
    __syncthreads();
    // x forward: n1 lines [y][kx]
    res = fft_in_lds<R>(other, res, ax0, n1, W0, R(-1));
    // the influence function, from its table ([kz][ky][kx])
    const R* gz = cp.infl + (size_t)kz * np;
    for (int idx = threadIdx.x; idx < np; idx += blockDim.x) {
        const R et = gz[idx];
        res[idx].x *= et;
        res[idx].y *= et;
    }
    __syncthreads();
    other = (res == A) ? B : A;
    // x backward
    res = fft_in_lds<R>(res, other, ax0, n1, W0, R(1));
    other = (res == A) ? B : A;
    // transpose [y][x] -> [x][y]
    for (int idx = threadIdx.x; idx < np; idx += blockDim.x) {
        const int y = idx / n0, x = idx - y * n0;
        other[x * n1 + y] = res[idx];
    }
    __syncthreads();
    // y backward
    res = fft_in_lds<R>(other, res, ax1, n0, W1, R(1));
    for (int idx = threadIdx.x; idx < np; idx += blockDim.x) base[idx] = R2<R>{res[idx].x, res[idx].y};
}

template <class R> static bool plane_small_pass(hipStream_t s, const FftPlan<R>& p, R2<R>* S, int n2h, int n2,
                                                ConvParams<R> cp) {
    const int n0 = p.ax[0].n, n1 = p.ax[1].n;
    const size_t lds = plane_small_lds<R>(n0, n1);
    if (lds > kPlaneSmallLds) return false;
    hipLaunchKernelGGL(k_fft_plane_small<R>, dim3(n2h), dim3(fft_threads<R>()), lds, s, S, p.ax[0], p.ax[1], p.tw[0],
                       p.tw[1], n2, cp);
    return true;
}

// ---------------------------------------------------------------------------
// FP32, n0 = n1 = 128: passes 2-4 fused, one workgroup per kz plane.
// The 128 x 128 complex plane S[kz][.][.] (128 KB) is held in LDS, and the
// y transform, the x transform, the influence function and both inverse
// transforms run on it without leaving the chip.  A line of 128 = 16 x 8 is
// transformed in two register steps (four-step algorithm): 16-point DFTs at
// stride 8 with the twiddles w^(b k1), then 8-point DFTs over contiguous
// groups of 8, which leaves X[k1 + 16 k2] in slot 8 k1 + k2 (digit-reversed
// order).  The influence function is applied in that order and the inverse
// runs the steps backwards, so no permutation pass is needed.
// LDS layout: element (x, y) at x * 152 + 9 (y / 8) + y % 8 (a pad word after
// every group of 8 and a row pitch of 8 mod 16 complex words keep the
// accesses of both steps, along rows and along columns, free of bank
// conflicts).
// ---------------------------------------------------------------------------
namespace plane {
constexpr int N = 128, A = 16, B = 8;
constexpr int kThreads = 1024;
constexpr int kPitch = (N / B) * (B + 1) + 8;  // 152
__device__ __forceinline__ int slot(int y) { return (y >> 3) * (B + 1) + (y & 7); }

typedef Cx<float> C;
// the register steps below serve both precisions (the z passes also run in FP64)
template <class R> __device__ __forceinline__ Cx<R> tw_s(const Cx<R>* tw, int j, R s) {
    return Cx<R>{tw[j].x, -s * tw[j].y};
}

// exp(s 2 pi i j / 16), j < 16
template <class R> __device__ __forceinline__ Cx<R> w16(int j, R s) {
    constexpr double c[16] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508979,
                              0.0, -0.38268343236508979, -0.70710678118654752, -0.92387953251128674,
                              -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508979,
                              0.0, 0.38268343236508979, 0.70710678118654752, 0.92387953251128674};
    constexpr double sn[16] = {0.0, 0.38268343236508979, 0.70710678118654752, 0.92387953251128674,
                               1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508979,
                               0.0, -0.38268343236508979, -0.70710678118654752, -0.92387953251128674,
                               -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508979};
    return Cx<R>{R(c[j]), s * R(sn[j])};
}
// natural-order DFTs: X[k] = sum_n v[n] exp(s 2 pi i n k / M)
template <class R> __device__ __forceinline__ void dft8(Cx<R>* v, R s) {
    Cx<R> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<R, 4>(e, s);
    dft_small<R, 4>(o, s);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const Cx<R> t = k == 0 ? o[0] : cmul(o[k], w16<R>(2 * k, s));
        v[k] = cadd(e[k], t);
        v[k + 4] = csub(e[k], t);
    }
}
template <class R> __device__ __forceinline__ void dft16(Cx<R>* v, R s) {
    Cx<R> d[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int m = 0; m < 4; m++) d[q][m] = v[4 * m + q];
        dft_small<R, 4>(d[q], s);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        Cx<R> y[4];
#pragma unroll
        for (int q = 0; q < 4; q++) y[q] = (q * k == 0) ? d[q][k] : cmul(d[q][k], w16<R>(q * k, s));
        dft_small<R, 4>(y, s);
#pragma unroll
        for (int l = 0; l < 4; l++) v[k + 4 * l] = y[l];
    }
}

// address of element k of line `line`: ROW lines run along y in row x = line,
// column lines along x in column y = line
template <bool ROW> __device__ __forceinline__ int addr(int line, int k) {
    return ROW ? line * kPitch + slot(k) : k * kPitch + slot(line);
}

// Step of stride 8 (16-point DFTs), one task per (line, b).  Forward: the
// result times w^(b k1) goes to slot 8 k1 + b.  Inverse (last step): natural
// order in, natural order out.
template <bool ROW, bool FWD> __device__ __forceinline__ void step16(C* P, const C* tw, float s) {
    const int t = threadIdx.x;
    const int line = ROW ? t >> 3 : t & 127, b = ROW ? t & 7 : t >> 7;
    C v[A];
#pragma unroll
    for (int a = 0; a < A; a++) v[a] = P[addr<ROW>(line, a * B + b)];
    dft16(v, s);
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
        C r = v[k1];
        if (FWD && b * k1 != 0) r = cmul(r, tw_s(tw, b * k1, s));
        P[addr<ROW>(line, k1 * B + b)] = r;
    }
}
// Step over contiguous groups (8-point DFTs), tasks (line, k1), two per
// thread.  Forward: slot 8 k1 + k2 receives X[k1 + 16 k2].  Inverse (first
// step): the inverse DFT over k2, times w^(-b k1), to slot 8 k1 + b.
template <bool ROW, bool FWD> __device__ __forceinline__ void step8(C* P, const C* tw, float s) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int t = threadIdx.x + h * kThreads;
        const int line = ROW ? t >> 4 : t & 127, k1 = ROW ? t & 15 : t >> 7;
        C v[B];
#pragma unroll
        for (int j = 0; j < B; j++) v[j] = P[addr<ROW>(line, k1 * B + j)];
        dft8(v, s);
#pragma unroll
        for (int j = 0; j < B; j++) {
            C r = v[j];
            if (!FWD && j * k1 != 0) r = cmul(r, tw_s(tw, j * k1, s));
            P[addr<ROW>(line, k1 * B + j)] = r;
        }
    }
}
}  // namespace plane

__global__ __launch_bounds__(plane::kThreads) void k_fft_plane128_convolve(R2<float>* __restrict__ S,
                                                                           const R2<float>* __restrict__ tw,
                                                                           int n2, ConvParams<float> cp) {
    if (skipped(cp.skip)) return;
    using namespace plane;
    __shared__ C P[N * kPitch];
    __shared__ C W[N];
    const int kz = blockIdx.x;
    R2<float>* base = S + (size_t)kz * N * N;
    if (threadIdx.x < N) W[threadIdx.x] = C{tw[threadIdx.x].x, tw[threadIdx.x].y};
    {
        R2<float> v[N * N / kThreads];
#pragma unroll
        for (int u = 0; u < N * N / kThreads; u++) v[u] = base[threadIdx.x + u * kThreads];
#pragma unroll
        for (int u = 0; u < N * N / kThreads; u++) {
            const int idx = threadIdx.x + u * kThreads;
            P[(idx >> 7) * kPitch + slot(idx & 127)] = C{v[u].x, v[u].y};
        }
    }
    __syncthreads();
    step16<true, true>(P, W, -1.f);
    __syncthreads();
    step8<true, true>(P, W, -1.f);
    __syncthreads();
    step16<false, true>(P, W, -1.f);
    __syncthreads();
    step8<false, true>(P, W, -1.f);
    __syncthreads();
    // the influence function at (kx, ky, kz) from its table ([kz][ky][kx]),
    // slots in digit-reversed order
    {
        const float* gz = cp.infl + (size_t)kz * N * N;
        for (int idx = threadIdx.x; idx < N * N; idx += kThreads) {
            const int sx = idx >> 7, sy = idx & 127;
            const int kx = (sx >> 3) + A * (sx & 7), ky = (sy >> 3) + A * (sy & 7);
            const float et = gz[ky * N + kx];
            C& v = P[sx * kPitch + slot(sy)];
            v.x *= et;
            v.y *= et;
        }
    }
    __syncthreads();
    step8<false, false>(P, W, 1.f);
    __syncthreads();
    step16<false, false>(P, W, 1.f);
    __syncthreads();
    step8<true, false>(P, W, 1.f);
    __syncthreads();
    step16<true, false>(P, W, 1.f);
    __syncthreads();
    for (int u = 0; u < N * N / kThreads; u++) {
        const int idx = threadIdx.x + u * kThreads;
        const C v = P[(idx >> 7) * kPitch + slot(idx & 127)];
        base[idx] = R2<float>{v.x, v.y};
    }
}

// ---------------------------------------------------------------------------
// FP32, n2 = 128: the z passes with the same register steps (16-point DFTs at
// stride 8, then 8-point DFTs over contiguous groups), 64 lines per
// workgroup.  Forward: the lines are staged in LDS with coalesced loads, the
// two steps run, and the second writes X[k1 + 16 k2] straight to the half
// spectrum S[kz][l] (kz <= 64), consecutive lines to consecutive words.
// Backward: the first inverse step reads S (the other half by Hermitian
// symmetry), the second leaves the real lines in LDS, and they are stored
// with coalesced writes.  LDS: line t at t * 145 + 9 (z / 8) + z % 8 (a
// pitch of 1 mod 16 complex words: 64 lines in one access spread over all
// banks).
// ---------------------------------------------------------------------------
namespace zline {
constexpr int N = 128, A = 16, B = 8;
// lines per workgroup: 64 in FP32, 32 in FP64 (74 KB of LDS either way)
template <class R> constexpr int lines() { return sizeof(R) == 4 ? 64 : 32; }
template <class R> constexpr int threads() { return lines<R>() * B; }
constexpr int kPitch = (N / B) * (B + 1) + 1;  // 145
__device__ __forceinline__ int at(int t, int z) { return t * kPitch + (z >> 3) * (B + 1) + (z & 7); }
}  // namespace zline

template <class R>
__global__ __launch_bounds__(zline::threads<R>()) void k_fft_z128_r2c(const R* __restrict__ Q, R2<R>* __restrict__ S,
                                                                     const R2<R>* __restrict__ tw, int nlines,
                                                                     const double* skip) {
    if (skipped(skip)) return;
    using namespace zline;
    typedef Cx<R> C;
    constexpr int T = lines<R>(), kThreads = threads<R>();
    constexpr int kV = 16 / (int)sizeof(R);  // reals per 16-byte load
    typedef R vec __attribute__((ext_vector_type(kV)));
    __shared__ C L[T * kPitch];
    __shared__ C W[N];
    const int l0 = blockIdx.x * T;
    if (threadIdx.x < N) W[threadIdx.x] = C{tw[threadIdx.x].x, tw[threadIdx.x].y};
    {
        // T lines of N reals in 16-byte loads
        constexpr int kLoads = T * N / kV / kThreads;
        const vec* src = reinterpret_cast<const vec*>(Q + (size_t)l0 * N);
        vec v[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; u++) v[u] = src[threadIdx.x + u * kThreads];
#pragma unroll
        for (int u = 0; u < kLoads; u++) {
            const int f = threadIdx.x + u * kThreads, t = f / (N / kV), z = (f % (N / kV)) * kV;
#pragma unroll
            for (int c = 0; c < kV; c++) L[at(t, z + c)] = C{v[u][c], R(0)};
        }
    }
    __syncthreads();
    {
        const int t = threadIdx.x % T, b = threadIdx.x / T;
        C v[A];
#pragma unroll
        for (int a = 0; a < A; a++) v[a] = L[at(t, a * B + b)];
        plane::dft16<R>(v, R(-1));
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) {
            C r = v[k1];
            if (b * k1 != 0) r = cmul(r, plane::tw_s<R>(W, b * k1, R(-1)));
            L[at(t, k1 * B + b)] = r;
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int task = threadIdx.x + h * kThreads;
        const int t = task % T, k1 = task / T;
        C v[B];
#pragma unroll
        for (int j = 0; j < B; j++) v[j] = L[at(t, k1 * B + j)];
        plane::dft8<R>(v, R(-1));
#pragma unroll
        for (int k2 = 0; k2 < B; k2++) {
            const int kz = k1 + A * k2;
            if (kz <= N / 2) S[(size_t)kz * nlines + l0 + t] = R2<R>{v[k2].x, v[k2].y};
        }
    }
}

template <class R>
__global__ __launch_bounds__(zline::threads<R>()) void k_fft_z128_c2r(R* __restrict__ Q, const R2<R>* __restrict__ S,
                                                                     const R2<R>* __restrict__ tw, int nlines,
                                                                     const double* skip) {
    if (skipped(skip)) return;
    using namespace zline;
    typedef Cx<R> C;
    constexpr int T = lines<R>(), kThreads = threads<R>();
    constexpr int kV = 16 / (int)sizeof(R);
    typedef R vec __attribute__((ext_vector_type(kV)));
    __shared__ C L[T * kPitch];
    __shared__ C W[N];
    const int l0 = blockIdx.x * T;
    if (threadIdx.x < N) W[threadIdx.x] = C{tw[threadIdx.x].x, tw[threadIdx.x].y};
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int task = threadIdx.x + h * kThreads;
        const int t = task % T, k1 = task / T;
        C v[B];
#pragma unroll
        for (int k2 = 0; k2 < B; k2++) {
            const int kz = k1 + A * k2;
            const bool cj = kz > N / 2;
            const R2<R> sv = S[(size_t)(cj ? N - kz : kz) * nlines + l0 + t];
            v[k2] = C{sv.x, cj ? -sv.y : sv.y};
        }
        plane::dft8<R>(v, R(1));
#pragma unroll
        for (int j = 0; j < B; j++) {
            C r = v[j];
            if (j * k1 != 0) r = cmul(r, plane::tw_s<R>(W, j * k1, R(1)));
            L[at(t, k1 * B + j)] = r;
        }
    }
    __syncthreads();
    {
        const int t = threadIdx.x % T, b = threadIdx.x / T;
        C v[A];
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) v[k1] = L[at(t, k1 * B + b)];
        plane::dft16<R>(v, R(1));
#pragma unroll
        for (int a = 0; a < A; a++) L[at(t, a * B + b)] = v[a];
    }
    __syncthreads();
    {
        constexpr int kStores = T * N / kV / kThreads;
        vec* dst = reinterpret_cast<vec*>(Q + (size_t)l0 * N);
#pragma unroll
        for (int u = 0; u < kStores; u++) {
            const int f = threadIdx.x + u * kThreads, t = f / (N / kV), z = (f % (N / kV)) * kV;
            vec o;
#pragma unroll
            for (int c = 0; c < kV; c++) o[c] = L[at(t, z + c)].x;
            dst[f] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
FastDiv make_fastdiv(unsigned d) {
    FastDiv f;
    f.d = d;
    f.m = ((1ull << 40) + d - 1) / d;  // ceil(2^40 / d): exact for x < 2^40 / d, here x < 2^20
    return f;
}

bool make_fft_axis(int n, FftAxis& ax) {
    int rad[16], c = 0, m = n;
    while (m % 4 == 0) { rad[c++] = 4; m /= 4; }
    while (m % 2 == 0) { rad[c++] = 2; m /= 2; }
    while (m % 3 == 0) { rad[c++] = 3; m /= 3; }
    while (m % 5 == 0) { rad[c++] = 5; m /= 5; }
    if (m != 1 || c > kFftMaxStages) return false;
    ax.n = n;
    ax.nst = c;
    ax.nd = make_fastdiv((unsigned)n);
    int Ns = 1;
    for (int k = 0; k < c; k++) {
        ax.st[k].rad = rad[k];
        ax.st[k].nb = make_fastdiv((unsigned)(n / rad[k]));
        ax.st[k].ns = make_fastdiv((unsigned)Ns);
        ax.st[k].twstep = n / (Ns * rad[k]);
        Ns *= rad[k];
    }
    return true;
}

// lines per workgroup: a power of two up to 16, two line buffers within
// kFftLdsBytes (one line when even that exceeds it: n up to 1024 needs 32 KB
// in FP64)
template <class R> static int lines_per_group(int n) {
    int T = 16;
    while (T > 1 && 2 * (size_t)T * n * sizeof(R2<R>) > (size_t)kFftLdsBytes) T >>= 1;
    return T;
}
static int log2i(int v) {
    int s = 0;
    while ((1 << s) < v) s++;
    return s;
}
template <class R> static size_t fft_lds(int T, int n) { return (2 * (size_t)T * n + n) * sizeof(R2<R>); }

// the z passes with the register-step kernels where they apply (n2 = 128,
// whole batches of 64 lines in FP32, 32 in FP64)
template <class R>
static bool zline_pass(hipStream_t s, const FftPlan<R>& p, R* Q, R2<R>* S, int dir, const double* skip) {
    constexpr int T = zline::lines<R>();
    const int nlines = p.ax[0].n * p.ax[1].n;
    if (p.ax[2].n != zline::N || nlines % T != 0) return false;
    if (dir < 0)
        hipLaunchKernelGGL(k_fft_z128_r2c<R>, dim3(nlines / T), dim3(zline::threads<R>()), 0, s, Q, S, p.tw[2], nlines,
                           skip);
    else
        hipLaunchKernelGGL(k_fft_z128_c2r<R>, dim3(nlines / T), dim3(zline::threads<R>()), 0, s, Q, S, p.tw[2], nlines,
                           skip);
    return true;
}

// passes 2-4 as one plane pass where that kernel applies (FP32, n0 = n1 = 128)
template <class R> static bool plane_pass(hipStream_t, const FftPlan<R>&, R2<R>*, int, int, ConvParams<R>) {
    return false;
}
template <>
bool plane_pass<float>(hipStream_t s, const FftPlan<float>& p, R2<float>* S, int n2h, int n2, ConvParams<float> cp) {
    if (p.ax[0].n != plane::N || p.ax[1].n != plane::N) return false;
    hipLaunchKernelGGL(k_fft_plane128_convolve, dim3(n2h), dim3(plane::kThreads), 0, s, S, p.tw[0], n2, cp);
    return true;
}

// The influence function of performMPIDReciprocalConvolution
// (MPIDReferenceForce.cpp:3329-3366), exp(-pi^2 m^2 / alpha^2) /
// (pi V m^2 |b(k)|^2) with m = box^-1 k, k folded to (-n/2, n/2], at every
// point of the half spectrum, in FP64.  The reference multiplies the full
// complex spectrum by it and keeps the real part of the inverse transform,
// which is the same as multiplying by its symmetric part
// G_s(k) = (G(k) + G(-k)) / 2.  G(-k) = G(k) except where k sits on a
// Nyquist plane of a triclinic cell (the fold keeps +n/2, so m(-k) != -m(k))
// and on the planes k = (n - 1) / 2 of an odd grid (the moduli's optimal-zeta
// factor folds there asymmetrically, :2780-2806); the half-spectrum
// transform needs the symmetric form, so the table holds G_s.
__device__ double influence_at(int kx, int ky, int kz, const InflParams& ip, const double* __restrict__ mod) {
    if ((kx | ky | kz) == 0) return 0.0;
    const int mx = kx < (ip.n[0] + 1) / 2 ? kx : kx - ip.n[0];
    const int my = ky < (ip.n[1] + 1) / 2 ? ky : ky - ip.n[1];
    const int mz = kz < (ip.n[2] + 1) / 2 ? kz : kz - ip.n[2];
    const double hx = mx * ip.r00, hy = mx * ip.r10 + my * ip.r11, hz = mx * ip.r20 + my * ip.r21 + mz * ip.r22;
    const double m2 = hx * hx + hy * hy + hz * hz;
    const double den = m2 * mod[kx] * mod[ip.n[0] + ky] * mod[ip.n[0] + ip.n[1] + kz];
    return ip.sc * exp(-ip.expf * m2) / den;
}

template <class R>
__global__ void k_influence_table(InflParams ip, const double* __restrict__ mod, R* __restrict__ out) {
    const int n0 = ip.n[0], n1 = ip.n[1], n2 = ip.n[2], n2h = n2 / 2 + 1;
    const size_t total = (size_t)n2h * n1 * n0;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(idx % n0), ky = (int)((idx / n0) % n1), kz = (int)(idx / ((size_t)n0 * n1));
        const double g = influence_at(kx, ky, kz, ip, mod);
        const double gm = influence_at((n0 - kx) % n0, (n1 - ky) % n1, (n2 - kz) % n2, ip, mod);
        out[idx] = (R)(0.5 * (g + gm));
    }
}

template <class R> void launch_influence_table(hipStream_t s, InflParams ip, const double* moduli64, R* out) {
    const size_t total = (size_t)(ip.n[2] / 2 + 1) * ip.n[1] * ip.n[0];
    const int blocks = (int)std::min<size_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_influence_table<R>, dim3(blocks), dim3(256), 0, s, ip, moduli64, out);
}
template void launch_influence_table<float>(hipStream_t, InflParams, const double*, float*);
template void launch_influence_table<double>(hipStream_t, InflParams, const double*, double*);

template <class R>
void launch_recip_convolution(hipStream_t s, const FftPlan<R>& p, R* Q, R2<R>* S, ConvParams<R> cp) {
    const int n0 = p.ax[0].n, n1 = p.ax[1].n, n2 = p.ax[2].n;
    const int n2h = n2 / 2 + 1;
    const int nlines = n0 * n1;
    const int Tz = lines_per_group<R>(n2);
    const int Ty = lines_per_group<R>(n1), nxb = (n0 + Ty - 1) / Ty;
    const int Tx = lines_per_group<R>(n0), nyb = (n1 + Tx - 1) / Tx;
    // 1. z forward
    if (!zline_pass<R>(s, p, Q, S, -1, cp.skip))
        hipLaunchKernelGGL(k_fft_z<R>, dim3((nlines + Tz - 1) / Tz), dim3(fft_threads<R>()), fft_lds<R>(Tz, n2), s, Q,
                           S, p.ax[2], Tz, log2i(Tz), p.tw[2], nlines, -1, cp.skip);
    if (!plane_pass<R>(s, p, S, n2h, n2, cp) && !plane_small_pass<R>(s, p, S, n2h, n2, cp)) {
        // 2. y forward
        hipLaunchKernelGGL(k_fft_y<R>, dim3(n2h * nxb), dim3(fft_threads<R>()), fft_lds<R>(Ty, n1), s, S, p.ax[1], Ty,
                           p.tw[1], n0, nxb, -1, cp.skip);
        // 3. x forward, convolution, x backward
        hipLaunchKernelGGL(k_fft_x_convolve<R>, dim3(n2h * nyb), dim3(fft_threads<R>()), fft_lds<R>(Tx, n0), s, S,
                           p.ax[0], Tx, log2i(Tx), p.tw[0], n1, n2, nyb, cp);
        // 4. y backward
        hipLaunchKernelGGL(k_fft_y<R>, dim3(n2h * nxb), dim3(fft_threads<R>()), fft_lds<R>(Ty, n1), s, S, p.ax[1], Ty,
                           p.tw[1], n0, nxb, 1, cp.skip);
    }
    // 5. z backward
    if (!zline_pass<R>(s, p, Q, S, 1, cp.skip))
        hipLaunchKernelGGL(k_fft_z<R>, dim3((nlines + Tz - 1) / Tz), dim3(fft_threads<R>()), fft_lds<R>(Tz, n2), s, Q,
                           S, p.ax[2], Tz, log2i(Tz), p.tw[2], nlines, 1, cp.skip);
}

template void launch_recip_convolution<float>(hipStream_t, const FftPlan<float>&, float*, R2<float>*,
                                              ConvParams<float>);
template void launch_recip_convolution<double>(hipStream_t, const FftPlan<double>&, double*, R2<double>*,
                                               ConvParams<double>);

}  // namespace mpid
