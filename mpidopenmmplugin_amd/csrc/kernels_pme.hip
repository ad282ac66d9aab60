// Reciprocal-space (PME) kernels of the MPIDForce path on gfx950: order-6
// B-spline spreading of multipoles and induced dipoles onto the real charge
// grid, and the gathering of the potential and its fractional derivatives
// (the FFT and convolution are in kernels_fft.hip).
//
// Follows MPIDReferencePmeForce (platforms/reference/src/SimTKReference/
// MPIDReferenceForce.cpp): computeBSplinePoint :2956-3044, computeMPIDBsplines
// :3049-3075, transformMultipolesToFractionalCoordinates :3077-3169,
// spreadFixedMultipolesOntoGrid :3269-3327, spreadInducedDipolesOnGrid
// :3532-3573, computeFixedPotentialFromGrid :3368-3530.  The grid has x
// slowest and z fastest, as the reference's; it holds only the real part,
// which is all the reference ever fills.
#include "engine.h"

#include <algorithm>

namespace mpid {

// Order-6 B-spline weights of the 6 grid points and their first ND derivatives:
// Cox-de Boor recursion, derivatives as backward differences of lower orders.
template <class R, int ND>
__device__ __forceinline__ void bspline6(R w, R th[kOrder][ND + 1]) {
    R c[kOrder + 1][kOrder + 1];
    c[1][0] = R(1);
#pragma unroll
    for (int p = 2; p <= kOrder; p++) {
#pragma unroll
        for (int m = 0; m < p; m++) {
            R left = (m >= 1) ? c[p - 1][m - 1] : R(0);
            R right = (m <= p - 2) ? c[p - 1][m] : R(0);
            c[p][m] = ((w + R(p - 1 - m)) * left + (R(m + 1) - w) * right) * R(1.0 / (p - 1));
        }
    }
#pragma unroll
    for (int r = 0; r <= ND; r++) {
        R cur[kOrder + 1];
        const int p = kOrder - r;
#pragma unroll
        for (int m = 0; m < kOrder + 1; m++) cur[m] = (m < p) ? c[p][m] : R(0);
        int len = p;
#pragma unroll
        for (int s = 0; s < r; s++) {
            R nxt[kOrder + 1];
#pragma unroll
            for (int m = 0; m < kOrder + 1; m++)
                nxt[m] = ((m >= 1 && m - 1 < len) ? cur[m - 1] : R(0)) - ((m < len) ? cur[m] : R(0));
            len++;
#pragma unroll
            for (int m = 0; m < kOrder + 1; m++) cur[m] = nxt[m];
        }
#pragma unroll
        for (int m = 0; m < kOrder; m++) th[m][r] = cur[m];
    }
}

__device__ __forceinline__ unsigned spread_bits_pme(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Grid origin and spline offset of one atom along an axis of n points from
// its fractional coordinate f along that box vector, as the reference
// computes them (f wrapped to [-1/2, 1/2), then shifted by one half box,
// MPIDReferenceForce.cpp:3049-3075).
template <class R>
__device__ __forceinline__ void grid_coord(R f, int n, int& ig, R& w) {
    f -= floor(f + R(0.5));
    const R fr = R(n) * (f + R(0.5));
    const int ifr = (int)floor(fr);
    w = fr - R(ifr);
    ig = ifr - kOrder + 1;
    ig %= n;
    if (ig < 0) ig += n;
}

// ---------------------------------------------------------------------------
// spreading
// ---------------------------------------------------------------------------
// Untiled spreading, for grids the tiled kernels do not take (sizes with no
// power-of-two tiling, or windows too large for LDS).  One wavefront per
// atom, its 64 lanes over the 6 x 6 x 6 stencil points (216 = 3 x 64 + 24):
// the adds of one wave go to distinct grid points, and every atom is a wave
// of its own, so that a small system still spreads over many CUs.  Every
// lane evaluates the atom's splines and selects its (ix, iy, iz) entries with
// unrolled compares (no dynamically indexed register arrays).
template <class R, int ND>
__device__ __forceinline__ void spline_at(const R (&t)[kOrder][ND + 1], int k, R (&out)[ND + 1]) {
#pragma unroll
    for (int d = 0; d <= ND; d++) out[d] = t[0][d];
#pragma unroll
    for (int m = 1; m < kOrder; m++)
        if (k == m) {
#pragma unroll
            for (int d = 0; d <= ND; d++) out[d] = t[m][d];
        }
}

constexpr int kSpreadWaves = 4;  // atoms (wavefronts) per 256-thread block

template <class R>
__global__ __launch_bounds__(256) void k_spread_fixed(int n, const R4<R>* __restrict__ pos, const R* __restrict__ lab,
                                                       PmeGeom<R> g, R* __restrict__ grid) {
    const int i = blockIdx.x * kSpreadWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const R4<R> p = pos[i];
    int ig[3];
    R w[3];
    {
        const V3<R> f = to_frac(mk(p.x, p.y, p.z), g.box);
        grid_coord(f.x, g.n[0], ig[0], w[0]);
        grid_coord(f.y, g.n[1], ig[1], w[1]);
        grid_coord(f.z, g.n[2], ig[2], w[2]);
    }
    R tx[kOrder][4], ty[kOrder][4], tz[kOrder][4];
    bspline6<R, 3>(w[0], tx);
    bspline6<R, 3>(w[1], ty);
    bspline6<R, 3>(w[2], tz);
    // fractional multipoles
    R mf[kMP];
    frac_moments(lab + (size_t)i * kMP, g.A, mf);
    const R q = mf[0];
    const R dx = mf[1], dy = mf[2], dz = mf[3];
    const R qxx = mf[4], qxy = mf[5], qxz = mf[6], qyy = mf[7], qyz = mf[8], qzz = mf[9];
    const R oxxx = mf[10], oxxy = mf[11], oxxz = mf[12], oxyy = mf[13], oxyz = mf[14], oxzz = mf[15];
    const R oyyy = mf[16], oyyz = mf[17], oyzz = mf[18], ozzz = mf[19];
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    for (int s = lane; s < kOrder * kOrder * kOrder; s += 64) {
        const int ix = s / (kOrder * kOrder), rem = s - ix * kOrder * kOrder;
        const int iy = rem / kOrder, iz = rem - iy * kOrder;
        R a[4], u[4], v[4];
        spline_at<R, 3>(tx, ix, a);
        spline_at<R, 3>(ty, iy, u);
        spline_at<R, 3>(tz, iz, v);
        const R t0 = q * u[0] * v[0] + dy * u[1] * v[0] + dz * u[0] * v[1] + qyy * u[2] * v[0] + qzz * u[0] * v[2] +
                     qyz * u[1] * v[1] + oyyy * u[3] * v[0] + oyyz * u[2] * v[1] + oyzz * u[1] * v[2] +
                     ozzz * u[0] * v[3];
        const R t1 = dx * u[0] * v[0] + qxy * u[1] * v[0] + qxz * u[0] * v[1] + oxyy * u[2] * v[0] +
                     oxyz * u[1] * v[1] + oxzz * u[0] * v[2];
        const R t2 = qxx * u[0] * v[0] + oxxy * u[1] * v[0] + oxxz * u[0] * v[1];
        const R t3 = oxxx * u[0] * v[0];
        int x = ig[0] + ix, y = ig[1] + iy, z = ig[2] + iz;
        if (x >= nx) x -= nx;
        if (y >= ny) y -= ny;
        if (z >= nz) z -= nz;
        atomicAdd(&grid[((size_t)x * ny + y) * nz + z], t0 * a[0] + t1 * a[1] + t2 * a[2] + t3 * a[3]);
    }
}

template <class R>
__global__ __launch_bounds__(256) void k_spread_induced(int n, const R4<R>* __restrict__ pos, const R* __restrict__ mu,
                                                         PmeGeom<R> g, R* __restrict__ grid) {
    const int i = blockIdx.x * kSpreadWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const V3<R> uf = frac_vector(mk(mu[3 * (size_t)i], mu[3 * (size_t)i + 1], mu[3 * (size_t)i + 2]), g.A);
    const R ux = uf.x, uy = uf.y, uz = uf.z;
    // a site without a dipole adds nothing (the whole wave leaves together)
    if (ux == R(0) && uy == R(0) && uz == R(0)) return;
    const R4<R> p = pos[i];
    int ig[3];
    R w[3];
    {
        const V3<R> f = to_frac(mk(p.x, p.y, p.z), g.box);
        grid_coord(f.x, g.n[0], ig[0], w[0]);
        grid_coord(f.y, g.n[1], ig[1], w[1]);
        grid_coord(f.z, g.n[2], ig[2], w[2]);
    }
    R tx[kOrder][2], ty[kOrder][2], tz[kOrder][2];
    bspline6<R, 1>(w[0], tx);
    bspline6<R, 1>(w[1], ty);
    bspline6<R, 1>(w[2], tz);
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    for (int s = lane; s < kOrder * kOrder * kOrder; s += 64) {
        const int ix = s / (kOrder * kOrder), rem = s - ix * kOrder * kOrder;
        const int iy = rem / kOrder, iz = rem - iy * kOrder;
        R a[2], u[2], v[2];
        spline_at<R, 1>(tx, ix, a);
        spline_at<R, 1>(ty, iy, u);
        spline_at<R, 1>(tz, iz, v);
        const R t0 = uy * u[1] * v[0] + uz * u[0] * v[1];
        const R t1 = ux * u[0] * v[0];
        int x = ig[0] + ix, y = ig[1] + iy, z = ig[2] + iz;
        if (x >= nx) x -= nx;
        if (y >= ny) y -= ny;
        if (z >= nz) z -= nz;
        atomicAdd(&grid[((size_t)x * ny + y) * nz + z], t0 * a[0] + t1 * a[1]);
    }
}

// ---------------------------------------------------------------------------
// Tiled spreading without global atomics.
//
// The atoms are sorted along a Morton curve of 1024^3 cells, so the atoms
// inside one cell of the coarser 2^m-per-axis Morton level are a contiguous
// range.  With n_k = 2^m * T_k, such a cell spans T_k grid spacings and the
// stencils of its atoms fall in a window of W_k <= T_k + 8 points that starts
// at c_k T_k + n_k/2 - 6 (the reference places atom x at n x / L + n / 2,
// MPIDReferenceForce.cpp:3061-3062; one point of slack on each side absorbs
// the rounding of the Morton key).  One workgroup per cell spreads its atoms
// into an LDS copy of the window, then stores the window to scratch; a fold
// pass adds, for every grid point, the windows that cover it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned compact_bits(unsigned v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x030000ffu;
    v = (v | (v >> 16)) & 0x000003ffu;
    return v;
}

// Range [tstart[t], tend[t]) of the atoms of Morton cell t.  With idx, over
// the subsequence idx[0 .. *count) of the sorted atoms (still in Morton order).
__global__ void k_tile_ranges(int n, const unsigned* __restrict__ keys, const int* __restrict__ idx,
                              const int* __restrict__ count, int shift, int* tstart, int* tend) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = count ? *count : n;
    if (s >= m) return;
    auto key = [&](int k) { return keys[idx ? idx[k] : k] >> shift; };
    const unsigned t = key(s);
    if (s == 0 || key(s - 1) != t) tstart[t] = s;
    if (s == m - 1 || key(s + 1) != t) tend[t] = s + 1;
}

template <class R, int MODE>
__global__ __launch_bounds__(256) void k_spread_tiled(const R4<R>* __restrict__ pos, const R* __restrict__ src,
                                                       PmeGeom<R> g, TileGeom tg, const int* __restrict__ atoms,
                                                       const int* __restrict__ tstart,
                                                       const int* __restrict__ tend, R* __restrict__ scratch,
                                                       R* __restrict__ grid) {
    constexpr int ND = MODE == 0 ? 3 : 1;   // spline derivatives needed
    constexpr int NC = MODE == 0 ? 20 : 3;  // fractional components
    constexpr int CH = 32;                  // atoms staged per pass
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int W0 = tg.W[0], W1 = tg.W[1], W2 = tg.W[2];
    const int WV = W0 * W1 * W2;
    // The window accumulates in FP64 for both precisions: on gfx950 LDS
    // atomic adds of doubles (ds_add_f64) keep up, those of floats stall.
    double* win = reinterpret_cast<double*>(smem_raw);
    R* sth = reinterpret_cast<R*>(win + WV);          // [CH][3][6][ND+1]
    R* smf = sth + CH * 3 * kOrder * (ND + 1);        // [CH][NC]
    int* sig = reinterpret_cast<int*>(smf + CH * NC); // [CH][3] local window origin of the stencil
    const int t = blockIdx.x;
    const int c[3] = {(int)compact_bits((unsigned)t >> 2), (int)compact_bits((unsigned)t >> 1),
                      (int)compact_bits((unsigned)t)};
    int w0[3];
#pragma unroll
    for (int k = 0; k < 3; k++) w0[k] = ((c[k] * tg.T[k] + (g.n[k] >> 1) - 6) % g.n[k] + g.n[k]) % g.n[k];
    for (int i = threadIdx.x; i < WV; i += blockDim.x) win[i] = 0.0;
    const int a_begin = tstart[t], a_end = tend[t];
    __syncthreads();
    for (int a0 = a_begin; a0 < a_end; a0 += CH) {
        const int cnt = min(CH, a_end - a0);
        if ((int)threadIdx.x < cnt) {
            const int a = atoms ? atoms[a0 + threadIdx.x] : a0 + threadIdx.x;
            const R4<R> p = pos[a];
            const V3<R> fp = to_frac(mk(p.x, p.y, p.z), g.box);
            const R px[3] = {fp.x, fp.y, fp.z};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                int ig;
                R w;
                grid_coord(px[k], g.n[k], ig, w);
                R th[kOrder][ND + 1];
                bspline6<R, ND>(w, th);
                for (int m = 0; m < kOrder; m++)
                    for (int r = 0; r <= ND; r++) sth[((threadIdx.x * 3 + k) * kOrder + m) * (ND + 1) + r] = th[m][r];
                sig[threadIdx.x * 3 + k] = (ig - w0[k] + g.n[k]) % g.n[k];
            }
            R* mf = smf + threadIdx.x * NC;
            if (MODE == 0) {
                frac_moments(src + (size_t)a * kMP, g.A, mf);
            } else {
                const V3<R> uf =
                    frac_vector(mk(src[3 * (size_t)a], src[3 * (size_t)a + 1], src[3 * (size_t)a + 2]), g.A);
                mf[0] = uf.x;
                mf[1] = uf.y;
                mf[2] = uf.z;
            }
        }
        __syncthreads();
        // one item = (atom, iy, iz): 6 points along x
        for (int it = threadIdx.x; it < cnt * kOrder * kOrder; it += blockDim.x) {
            const int la = it / (kOrder * kOrder);
            const int iy = (it / kOrder) % kOrder, iz = it % kOrder;
            const R* tx = sth + ((la * 3 + 0) * kOrder) * (ND + 1);
            const R* ty = sth + ((la * 3 + 1) * kOrder + iy) * (ND + 1);
            const R* tz = sth + ((la * 3 + 2) * kOrder + iz) * (ND + 1);
            const R* mf = smf + la * NC;
            R t0, t1, t2 = R(0), t3 = R(0);
            if (MODE == 0) {
                const R u0 = ty[0], u1 = ty[1], u2 = ty[2], u3 = ty[ND >= 3 ? 3 : 0];
                const R v0 = tz[0], v1 = tz[1], v2 = tz[2], v3 = tz[ND >= 3 ? 3 : 0];
                t0 = mf[0] * u0 * v0 + mf[2] * u1 * v0 + mf[3] * u0 * v1 + mf[7] * u2 * v0 + mf[9] * u0 * v2 +
                     mf[8] * u1 * v1 + mf[16] * u3 * v0 + mf[17] * u2 * v1 + mf[18] * u1 * v2 + mf[19] * u0 * v3;
                t1 = mf[1] * u0 * v0 + mf[5] * u1 * v0 + mf[6] * u0 * v1 + mf[13] * u2 * v0 + mf[14] * u1 * v1 +
                     mf[15] * u0 * v2;
                t2 = mf[4] * u0 * v0 + mf[11] * u1 * v0 + mf[12] * u0 * v1;
                t3 = mf[10] * u0 * v0;
            } else {
                t0 = mf[1] * ty[1] * tz[0] + mf[2] * ty[0] * tz[1];
                t1 = mf[0] * ty[0] * tz[0];
            }
            const int ly = sig[la * 3 + 1] + iy, lz = sig[la * 3 + 2] + iz;
            const int lx0 = sig[la * 3 + 0];
            const bool inside_yz = ly < W1 && lz < W2;
#pragma unroll
            for (int ix = 0; ix < kOrder; ix++) {
                R v = t0 * tx[ix * (ND + 1)] + t1 * tx[ix * (ND + 1) + 1];
                if (MODE == 0) v += t2 * tx[ix * (ND + 1) + 2] + t3 * tx[ix * (ND + 1) + ND];
                const int lx = lx0 + ix;
                if (inside_yz && lx < W0) {
                    atomicAdd(&win[(lx * W1 + ly) * W2 + lz], (double)v);
                } else {
                    // outside the window (not expected): straight to the grid
                    const int x = (w0[0] + lx) % g.n[0], y = (w0[1] + ly) % g.n[1], z = (w0[2] + lz) % g.n[2];
                    atomicAdd(&grid[((size_t)x * g.n[1] + y) * g.n[2] + z], v);
                }
            }
        }
        __syncthreads();
    }
    R* out = scratch + (size_t)t * WV;
    for (int i = threadIdx.x; i < WV; i += blockDim.x) out[i] = (R)win[i];
}

// ---------------------------------------------------------------------------
// Owner-computes spreading (tiles of at least 8 points per axis).
// The grid points are partitioned into cores: core c is the T0 x T1 x T2 block
// starting at (c T + n/2 - 6) mod n, the window origin of cell c above.  An
// atom of cell c (cell of its Morton key, from the FP64 position) has
// fractional grid coordinate fr in [c T + n/2, (c + 1) T + n/2) mod n, so its
// stencil start ig = floor(fr) - 5 lies in [c T + n/2 - 5, (c + 1) T + n/2 - 6]
// up to one point of rounding, i.e. in [origin(c), origin(c + 1)]: it meets
// the cores c and c + 1 only (T >= 8).  The workgroup of core c therefore
// scans the cells c - 1 and c along each axis (8 cells), keeps the atoms
// whose stencil meets the core, adds their contributions to an FP64 LDS
// copy of the core and writes the core to the grid once.  No scratch, no
// fold and no clearing of the grid: every point is written by its owner.
// ---------------------------------------------------------------------------
// candidate atoms examined per pass (the slots of their spline weights in LDS)
template <class R, int MODE> constexpr int owner_cand() { return MODE == 0 ? 64 : 128; }
// stencil rows (of 6) along an axis inside a core of T points, for a stencil
// starting at r relative to the core (r in [-5, T - 1] for a stencil that meets it)
__device__ __forceinline__ int owner_rows(int r, int T) { return min(kOrder, T - r) - max(0, -r); }

template <class R, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_spread_owner(const R4<R>* __restrict__ pos, const R* __restrict__ src,
                                                      PmeGeom<R> g, TileGeom tg, const int* __restrict__ atoms,
                                                      const int* __restrict__ tstart, const int* __restrict__ tend,
                                                      R* __restrict__ grid) {
    if (skipped(g.skip)) return;
    constexpr int ND = MODE == 0 ? 3 : 1;   // spline derivatives needed
    constexpr int NC = MODE == 0 ? 20 : 3;  // fractional components
    constexpr int kOwnerCand = owner_cand<R, MODE>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int T0 = tg.T[0], T1 = tg.T[1], T2 = tg.T[2];
    const int TV = T0 * T1 * T2;
    double* core = reinterpret_cast<double*>(smem_raw);
    R* sth = reinterpret_cast<R*>(core + TV);                       // [kOwnerCand][3][6][ND+1]
    R* smf = sth + kOwnerCand * 3 * kOrder * (ND + 1);              // [kOwnerCand][NC]
    int* srel = reinterpret_cast<int*>(smf + kOwnerCand * NC);      // [kOwnerCand][3] stencil start - core origin
    __shared__ int sbeg[8], send_[8], soff[9], snv, snrows;
    __shared__ int sfirst[kOwnerCand];
    static_assert(kOwnerCand % kTile == 0 && kOwnerCand <= 256, "slots: whole wavefronts, at most the block");
    const int t = blockIdx.x;
    const int nt = tg.tiles_per_axis;
    const int c[3] = {(int)compact_bits((unsigned)t >> 2), (int)compact_bits((unsigned)t >> 1),
                      (int)compact_bits((unsigned)t)};
    int w0[3];
#pragma unroll
    for (int k = 0; k < 3; k++) w0[k] = ((c[k] * tg.T[k] + (g.n[k] >> 1) - 6) % g.n[k] + g.n[k]) % g.n[k];
    for (int i = threadIdx.x; i < TV; i += blockDim.x) core[i] = 0.0;
    if (threadIdx.x < 8) {
        const int d = threadIdx.x;
        const unsigned tx = (unsigned)((c[0] - ((d >> 2) & 1) + nt) % nt),
                       ty = (unsigned)((c[1] - ((d >> 1) & 1) + nt) % nt), tz = (unsigned)((c[2] - (d & 1) + nt) % nt);
        const unsigned tile = (spread_bits_pme(tx) << 2) | (spread_bits_pme(ty) << 1) | spread_bits_pme(tz);
        sbeg[d] = tstart[tile];
        send_[d] = tend[tile];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // cells c - 1 and c coincide only with one cell per axis: count each cell once
        int o = 0;
        for (int d = 0; d < 8; d++) {
            bool dup = false;
            for (int e = 0; e < d; e++) dup = dup || (sbeg[e] == sbeg[d] && send_[e] == send_[d] && send_[d] > sbeg[d]);
            soff[d] = o;
            if (!dup) o += max(0, send_[d] - sbeg[d]);
            else send_[d] = sbeg[d];
        }
        soff[8] = o;
    }
    __syncthreads();
    const int ncand = soff[8];
    for (int c0 = 0; c0 < ncand; c0 += kOwnerCand) {
        if (threadIdx.x == 0) snv = 0;
        __syncthreads();
        if ((int)threadIdx.x < kOwnerCand && c0 + (int)threadIdx.x < ncand) {
            const int q = c0 + threadIdx.x;
            int d = 0;
            while (soff[d + 1] <= q) d++;
            const int s = sbeg[d] + (q - soff[d]);
            const int a = atoms ? atoms[s] : s;
            const R4<R> p = pos[a];
            const V3<R> fp = to_frac(mk(p.x, p.y, p.z), g.box);
            const R px[3] = {fp.x, fp.y, fp.z};
            int ig[3], rel[3];
            R w[3];
            bool meets = true;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                grid_coord(px[k], g.n[k], ig[k], w[k]);
                int r = ig[k] - w0[k];
                if (r < -(g.n[k] >> 1)) r += g.n[k];
                if (r >= g.n[k] - (g.n[k] >> 1)) r -= g.n[k];
                rel[k] = r;
                meets = meets && r + kOrder - 1 >= 0 && r < tg.T[k];
            }
            if (meets) {
                const int slot = atomicAdd(&snv, 1);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    R th[kOrder][ND + 1];
                    bspline6<R, ND>(w[k], th);
                    for (int m = 0; m < kOrder; m++)
                        for (int r = 0; r <= ND; r++) sth[((slot * 3 + k) * kOrder + m) * (ND + 1) + r] = th[m][r];
                    srel[slot * 3 + k] = rel[k];
                }
                R* mf = smf + slot * NC;
                    if (MODE == 0) {
                    frac_moments(src + (size_t)a * kMP, g.A, mf);
                } else {
                    const V3<R> uf =
                        frac_vector(mk(src[3 * (size_t)a], src[3 * (size_t)a + 1], src[3 * (size_t)a + 2]), g.A);
                    mf[0] = uf.x;
                    mf[1] = uf.y;
                    mf[2] = uf.z;
                }
            }
        }
        __syncthreads();
        const int nv = snv;
        // The listed atoms' stencil rows (iy, iz) inside the core, numbered
        // atom by atom (sfirst[la]: the first row of atom la), so that every
        // thread's item below is a row with points to add: about 14 of an
        // atom's 36 rows meet the core (T = 8).  One wavefront counts them.
        if (threadIdx.x < kTile) {
            int cnt[kOwnerCand / kTile], inc[kOwnerCand / kTile];
#pragma unroll
            for (int h = 0; h < kOwnerCand / kTile; h++) {
                const int la = h * kTile + (int)threadIdx.x;
                cnt[h] = la < nv ? owner_rows(srel[la * 3 + 1], T1) * owner_rows(srel[la * 3 + 2], T2) : 0;
                inc[h] = cnt[h];
            }
#pragma unroll
            for (int off = 1; off < kTile; off <<= 1) {
#pragma unroll
                for (int h = 0; h < kOwnerCand / kTile; h++) {
                    const int v = __shfl_up(inc[h], off);
                    if ((int)threadIdx.x >= off) inc[h] += v;
                }
            }
            int base = 0;
#pragma unroll
            for (int h = 0; h < kOwnerCand / kTile; h++) {
                const int la = h * kTile + (int)threadIdx.x;
                if (la < nv) sfirst[la] = base + inc[h] - cnt[h];
                base += __shfl(inc[h], kTile - 1);
            }
            if (threadIdx.x == 0) snrows = base;
        }
        __syncthreads();
        const int nrows = snrows;
        // one item = (atom, iy, iz): the points of the stencil row along x inside the core
        for (int it = threadIdx.x; it < nrows; it += blockDim.x) {
            // the atom: the last la with sfirst[la] <= it (every atom has a row)
            int la = 0;
#pragma unroll
            for (int step = kOwnerCand / 2; step >= 1; step >>= 1)
                if (la + step < nv && sfirst[la + step] <= it) la += step;
            const int ry = srel[la * 3 + 1], rz = srel[la * 3 + 2];
            const int y0 = max(0, -ry), z0 = max(0, -rz);
            const int nz = min(kOrder, T2 - rz) - z0;
            const int loc = it - sfirst[la], dy = loc / nz;
            const int iy = y0 + dy, iz = z0 + (loc - dy * nz);
            const int ly = ry + iy, lz = rz + iz;
            const R* tx = sth + ((la * 3 + 0) * kOrder) * (ND + 1);
            const R* ty = sth + ((la * 3 + 1) * kOrder + iy) * (ND + 1);
            const R* tz = sth + ((la * 3 + 2) * kOrder + iz) * (ND + 1);
            const R* mf = smf + la * NC;
            R t0, t1, t2 = R(0), t3 = R(0);
            if (MODE == 0) {
                const R u0 = ty[0], u1 = ty[1], u2 = ty[2], u3 = ty[ND >= 3 ? 3 : 0];
                const R v0 = tz[0], v1 = tz[1], v2 = tz[2], v3 = tz[ND >= 3 ? 3 : 0];
                t0 = mf[0] * u0 * v0 + mf[2] * u1 * v0 + mf[3] * u0 * v1 + mf[7] * u2 * v0 + mf[9] * u0 * v2 +
                     mf[8] * u1 * v1 + mf[16] * u3 * v0 + mf[17] * u2 * v1 + mf[18] * u1 * v2 + mf[19] * u0 * v3;
                t1 = mf[1] * u0 * v0 + mf[5] * u1 * v0 + mf[6] * u0 * v1 + mf[13] * u2 * v0 + mf[14] * u1 * v1 +
                     mf[15] * u0 * v2;
                t2 = mf[4] * u0 * v0 + mf[11] * u1 * v0 + mf[12] * u0 * v1;
                t3 = mf[10] * u0 * v0;
            } else {
                t0 = mf[1] * ty[1] * tz[0] + mf[2] * ty[0] * tz[1];
                t1 = mf[0] * ty[0] * tz[0];
            }
            const int lx0 = srel[la * 3 + 0];
#pragma unroll
            for (int ix = 0; ix < kOrder; ix++) {
                const int lx = lx0 + ix;
                if (lx < 0 || lx >= T0) continue;
                R v = t0 * tx[ix * (ND + 1)] + t1 * tx[ix * (ND + 1) + 1];
                if (MODE == 0) v += t2 * tx[ix * (ND + 1) + 2] + t3 * tx[ix * (ND + 1) + ND];
                atomicAdd(&core[(lx * T1 + ly) * T2 + lz], (double)v);
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < TV; i += blockDim.x) {
        const int lx = i / (T1 * T2), rem = i - lx * T1 * T2;
        const int ly = rem / T2, lz = rem - ly * T2;
        int gx = w0[0] + lx, gy = w0[1] + ly, gz = w0[2] + lz;
        if (gx >= g.n[0]) gx -= g.n[0];
        if (gy >= g.n[1]) gy -= g.n[1];
        if (gz >= g.n[2]) gz -= g.n[2];
        grid[((size_t)gx * g.n[1] + gy) * g.n[2] + gz] = (R)core[i];
    }
}

// grid(x,y,z) = sum of the windows covering the point (plus what is already there)
template <class R>
__global__ __launch_bounds__(256) void k_fold(PmeGeom<R> g, TileGeom tg, const R* __restrict__ scratch,
                                               R* __restrict__ grid) {
    const size_t total = (size_t)g.n[0] * g.n[1] * g.n[2];
    const int WV = tg.W[0] * tg.W[1] * tg.W[2];
    const int nt = tg.tiles_per_axis;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
        const int gz = (int)(id % g.n[2]);
        const int gy = (int)((id / g.n[2]) % g.n[1]);
        const int gx = (int)(id / ((size_t)g.n[1] * g.n[2]));
        const int gg[3] = {gx, gy, gz};
        int cc[3], rr[3], nj[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int q = ((gg[k] - (g.n[k] >> 1) + 6) % g.n[k] + g.n[k]) % g.n[k];
            cc[k] = q / tg.T[k];
            rr[k] = q - cc[k] * tg.T[k];
            nj[k] = min((tg.W[k] - 1 - rr[k]) / tg.T[k] + 1, nt);  // windows c, c-1, ... covering the point
        }
        // at most 3 windows per dimension cover a point (W <= T + 8, T >= 6);
        // the loop is unrolled so that all the loads are in flight together
        R sum = grid[id];
#pragma unroll
        for (int jx = 0; jx < 3; jx++)
#pragma unroll
            for (int jy = 0; jy < 3; jy++)
#pragma unroll
                for (int jz = 0; jz < 3; jz++) {
                    if (jx < nj[0] && jy < nj[1] && jz < nj[2]) {
                        const unsigned tx = (unsigned)((cc[0] - jx + nt) % nt),
                                       ty = (unsigned)((cc[1] - jy + nt) % nt),
                                       tz = (unsigned)((cc[2] - jz + nt) % nt);
                        const unsigned tile =
                            (spread_bits_pme(tx) << 2) | (spread_bits_pme(ty) << 1) | spread_bits_pme(tz);
                        const int lx = rr[0] + jx * tg.T[0], ly = rr[1] + jy * tg.T[1], lz = rr[2] + jz * tg.T[2];
                        sum += scratch[(size_t)tile * WV + (lx * tg.W[1] + ly) * tg.W[2] + lz];
                    }
                }
        grid[id] = sum;
    }
}

// When W = T + 8 <= 2T in every dimension, the point l (local to the core
// of its cell c) lies in window c at l and in window c - 1 at l + T when
// l < 8, and in no other window.  One workgroup per cell then sums the
// (at most 8) windows over its core, the neighbour cells' indices computed
// once per workgroup.
template <class R>
__global__ __launch_bounds__(256) void k_fold_tiles(PmeGeom<R> g, TileGeom tg, FastDiv t2, FastDiv t12,
                                                     const R* __restrict__ scratch, R* __restrict__ grid) {
    const int t = blockIdx.x;
    const int nt = tg.tiles_per_axis;
    const int c[3] = {(int)compact_bits((unsigned)t >> 2), (int)compact_bits((unsigned)t >> 1),
                      (int)compact_bits((unsigned)t)};
    int w0[3];
#pragma unroll
    for (int k = 0; k < 3; k++) w0[k] = ((c[k] * tg.T[k] + (g.n[k] >> 1) - 6) % g.n[k] + g.n[k]) % g.n[k];
    size_t base[8];
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const unsigned tx = (unsigned)((c[0] - ((d >> 2) & 1) + nt) % nt);
        const unsigned ty = (unsigned)((c[1] - ((d >> 1) & 1) + nt) % nt);
        const unsigned tz = (unsigned)((c[2] - (d & 1) + nt) % nt);
        const unsigned tile = (spread_bits_pme(tx) << 2) | (spread_bits_pme(ty) << 1) | spread_bits_pme(tz);
        base[d] = (size_t)tile * tg.W[0] * tg.W[1] * tg.W[2];
    }
    const int T0 = tg.T[0], T1 = tg.T[1], T2 = tg.T[2];
    const int W1 = tg.W[1], W2 = tg.W[2];
    const int ov0 = tg.W[0] - T0, ov1 = W1 - T1, ov2 = W2 - T2;  // = 8
    const int core = T0 * T1 * T2;
    for (int p = threadIdx.x; p < core; p += blockDim.x) {
        const int lx = (int)fdiv((unsigned)p, t12), rem = p - lx * T1 * T2;
        const int ly = (int)fdiv((unsigned)rem, t2), lz = rem - ly * T2;
        int gx = w0[0] + lx, gy = w0[1] + ly, gz = w0[2] + lz;
        if (gx >= g.n[0]) gx -= g.n[0];
        if (gy >= g.n[1]) gy -= g.n[1];
        if (gz >= g.n[2]) gz -= g.n[2];
        const size_t id = ((size_t)gx * g.n[1] + gy) * g.n[2] + gz;
        R sum = grid[id];
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const int dx = (d >> 2) & 1, dy = (d >> 1) & 1, dz = d & 1;
            if ((!dx || lx < ov0) && (!dy || ly < ov1) && (!dz || lz < ov2))
                sum += scratch[base[d] + ((size_t)(lx + dx * T0) * W1 + (ly + dy * T1)) * W2 + lz + dz * T2];
        }
        grid[id] = sum;
    }
}

// ---------------------------------------------------------------------------
// gathering
// ---------------------------------------------------------------------------
// Gathering: 8 lanes per atom, lane iz < 6 sums the 36 points of its z plane
// of the stencil (all 36 loads issued together), and the planes meet in a
// lane reduction.

// Fractional potential derivatives phi_abc, a+b+c <= 4, in the reference's
// order (MPIDReferenceForce.cpp:3498-3532).
template <class R>
__global__ __launch_bounds__(256) void k_gather_phi(int n, const R4<R>* __restrict__ pos, PmeGeom<R> g,
                                                     const R* __restrict__ grid, R* __restrict__ phi) {
    const int gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int t = gt >> 3, iz = gt & 7;
    const bool tv = t < n;
    const int i = tv ? t : n - 1;
    const R4<R> p = pos[i];
    int ig[3];
    R w[3];
    {
        const V3<R> f = to_frac(mk(p.x, p.y, p.z), g.box);
        grid_coord(f.x, g.n[0], ig[0], w[0]);
        grid_coord(f.y, g.n[1], ig[1], w[1]);
        grid_coord(f.z, g.n[2], ig[2], w[2]);
    }
    R tx[kOrder][5], ty[kOrder][5], tz[kOrder][5];
    bspline6<R, 4>(w[0], tx);
    bspline6<R, 4>(w[1], ty);
    bspline6<R, 4>(w[2], tz);
    const int ny = g.n[1], nz = g.n[2];
    R sv[35];
#pragma unroll
    for (int k = 0; k < 35; k++) sv[k] = R(0);
    if (tv && iz < kOrder) {
        int z = ig[2] + iz;
        if (z >= nz) z -= nz;
        // each row of six grid values is used as it arrives (holding all 36
        // put the FP64 kernel's arrays in scratch memory)
        R b00 = 0, b10 = 0, b01 = 0, b20 = 0, b11 = 0, b02 = 0, b30 = 0, b21 = 0, b12 = 0, b03 = 0, b40 = 0, b04 = 0,
          b31 = 0, b13 = 0, b22 = 0;
#pragma unroll
        for (int iy = 0; iy < kOrder; iy++) {
            int y = ig[1] + iy;
            if (y >= ny) y -= ny;
            R row[kOrder];
#pragma unroll
            for (int ix = 0; ix < kOrder; ix++) {
                int x = ig[0] + ix;
                if (x >= g.n[0]) x -= g.n[0];
                row[ix] = grid[((size_t)x * ny + y) * nz + z];
            }
            R a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
#pragma unroll
            for (int ix = 0; ix < kOrder; ix++) {
                const R v = row[ix];
                a0 += v * tx[ix][0];
                a1 += v * tx[ix][1];
                a2 += v * tx[ix][2];
                a3 += v * tx[ix][3];
                a4 += v * tx[ix][4];
            }
            const R u0 = ty[iy][0], u1 = ty[iy][1], u2 = ty[iy][2], u3 = ty[iy][3], u4 = ty[iy][4];
            b00 += a0 * u0; b10 += a1 * u0; b01 += a0 * u1; b20 += a2 * u0; b11 += a1 * u1;
            b02 += a0 * u2; b30 += a3 * u0; b21 += a2 * u1; b12 += a1 * u2; b03 += a0 * u3;
            b40 += a4 * u0; b04 += a0 * u4; b31 += a3 * u1; b13 += a1 * u3; b22 += a2 * u2;
        }
        // this lane's z weights (a select over the rows: a runtime index
        // into tz would put the table in scratch memory)
        R v0 = tz[0][0], v1 = tz[0][1], v2 = tz[0][2], v3 = tz[0][3], v4 = tz[0][4];
#pragma unroll
        for (int m = 1; m < kOrder; m++)
            if (iz == m) {
                v0 = tz[m][0];
                v1 = tz[m][1];
                v2 = tz[m][2];
                v3 = tz[m][3];
                v4 = tz[m][4];
            }
        // order of the reference's 35 derivatives (tuv000 ... tuv112)
        sv[0] = b00 * v0; sv[1] = b10 * v0; sv[2] = b01 * v0; sv[3] = b00 * v1;
        sv[4] = b20 * v0; sv[5] = b02 * v0; sv[6] = b00 * v2; sv[7] = b11 * v0; sv[8] = b10 * v1; sv[9] = b01 * v1;
        sv[10] = b30 * v0; sv[11] = b03 * v0; sv[12] = b00 * v3; sv[13] = b21 * v0; sv[14] = b20 * v1;
        sv[15] = b12 * v0; sv[16] = b02 * v1; sv[17] = b10 * v2; sv[18] = b01 * v2; sv[19] = b11 * v1;
        sv[20] = b40 * v0; sv[21] = b04 * v0; sv[22] = b00 * v4; sv[23] = b31 * v0; sv[24] = b30 * v1;
        sv[25] = b13 * v0; sv[26] = b03 * v1; sv[27] = b10 * v3; sv[28] = b01 * v3; sv[29] = b22 * v0;
        sv[30] = b20 * v2; sv[31] = b02 * v2; sv[32] = b21 * v1; sv[33] = b12 * v1; sv[34] = b11 * v2;
    }
#pragma unroll
    for (int k = 0; k < 35; k++) {
        sv[k] += __shfl_xor(sv[k], 1);
        sv[k] += __shfl_xor(sv[k], 2);
        sv[k] += __shfl_xor(sv[k], 4);
    }
    // every lane of the group holds the sums: the eight write the atom's 35
    // values together, eight consecutive words per store
    if (tv) {
        R* o = phi + (size_t)i * 35;
#pragma unroll
        for (int r = 0; r < 5; r++) {
            R v = sv[8 * r];
#pragma unroll
            for (int j = 1; j < 8; j++)
                if (8 * r + j < 35 && iz == j) v = sv[8 * r + j];
            if (iz + 8 * r < 35) o[iz + 8 * r] = v;
        }
    }
}

// Field -grad(phi) only (the matrix-vector product of the solver needs no more).
template <class R>
__global__ __launch_bounds__(256) void k_gather_field(int natoms, const int* __restrict__ atoms,
                                                       const R4<R>* __restrict__ pos, PmeGeom<R> g,
                                                       const R* __restrict__ grid, R* __restrict__ field) {
    if (skipped(g.skip)) return;
    const int gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int t = gt >> 3, iz = gt & 7;
    const bool tv = t < natoms;
    const int tc = tv ? t : natoms - 1;
    const int i = atoms ? atoms[tc] : tc;
    const R4<R> p = pos[i];
    int ig[3];
    R w[3];
    {
        const V3<R> f = to_frac(mk(p.x, p.y, p.z), g.box);
        grid_coord(f.x, g.n[0], ig[0], w[0]);
        grid_coord(f.y, g.n[1], ig[1], w[1]);
        grid_coord(f.z, g.n[2], ig[2], w[2]);
    }
    R tx[kOrder][2], ty[kOrder][2], tz[kOrder][2];
    bspline6<R, 1>(w[0], tx);
    bspline6<R, 1>(w[1], ty);
    bspline6<R, 1>(w[2], tz);
    const int ny = g.n[1], nz = g.n[2];
    R s100 = 0, s010 = 0, s001 = 0;
    if (tv && iz < kOrder) {
        int z = ig[2] + iz;
        if (z >= nz) z -= nz;
        R q[kOrder][kOrder];
#pragma unroll
        for (int iy = 0; iy < kOrder; iy++) {
            int y = ig[1] + iy;
            if (y >= ny) y -= ny;
#pragma unroll
            for (int ix = 0; ix < kOrder; ix++) {
                int x = ig[0] + ix;
                if (x >= g.n[0]) x -= g.n[0];
                q[iy][ix] = grid[((size_t)x * ny + y) * nz + z];
            }
        }
        R b00 = 0, b10 = 0, b01 = 0;
#pragma unroll
        for (int iy = 0; iy < kOrder; iy++) {
            R a0 = 0, a1 = 0;
#pragma unroll
            for (int ix = 0; ix < kOrder; ix++) {
                a0 += q[iy][ix] * tx[ix][0];
                a1 += q[iy][ix] * tx[ix][1];
            }
            b00 += a0 * ty[iy][0];
            b10 += a1 * ty[iy][0];
            b01 += a0 * ty[iy][1];
        }
        s100 = b10 * tz[iz][0];
        s010 = b01 * tz[iz][0];
        s001 = b00 * tz[iz][1];
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        s100 += __shfl_xor(s100, off);
        s010 += __shfl_xor(s010, off);
        s001 += __shfl_xor(s001, off);
    }
    if (tv && iz == 0) {
        const V3<R> e = cart_vector(mk(s100, s010, s001), g.A);
        field[3 * (size_t)i + 0] = -e.x;
        field[3 * (size_t)i + 1] = -e.y;
        field[3 * (size_t)i + 2] = -e.z;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class R>
void launch_spread_fixed(hipStream_t s, int n, const R4<R>* pos, const R* lab, PmeGeom<R> g, R* grid) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spread_fixed<R>, dim3((n + kSpreadWaves - 1) / kSpreadWaves), dim3(64 * kSpreadWaves), 0, s,
                       n, pos, lab, g, grid);
}
template <class R>
void launch_spread_induced(hipStream_t s, int n, const R4<R>* pos, const R* mu, PmeGeom<R> g, R* grid) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spread_induced<R>, dim3((n + kSpreadWaves - 1) / kSpreadWaves), dim3(64 * kSpreadWaves), 0,
                       s, n, pos, mu, g, grid);
}
void launch_tile_ranges(hipStream_t s, int n, const unsigned* keys_sorted, const int* idx, const int* count,
                        int shift, int* tstart, int* tend) {
    hipLaunchKernelGGL(k_tile_ranges, dim3((n + 255) / 256), dim3(256), 0, s, n, keys_sorted, idx, count, shift, tstart,
                       tend);
}

template <class R, int MODE> static size_t tiled_lds(const TileGeom& tg) {
    constexpr int ND = MODE == 0 ? 3 : 1;
    constexpr int NC = MODE == 0 ? 20 : 3;
    return (size_t)tg.W[0] * tg.W[1] * tg.W[2] * sizeof(double) + 32 * (3 * kOrder * (ND + 1) + NC) * sizeof(R) +
           32 * 3 * sizeof(int);
}
size_t spread_tiled_lds_bytes(const TileGeom& tg, int double_precision) {
    return double_precision ? tiled_lds<double, 0>(tg) : tiled_lds<float, 0>(tg);
}

template <class R>
void launch_spread_fixed_tiled(hipStream_t s, const R4<R>* pos, const R* lab, PmeGeom<R> g, TileGeom tg,
                               const int* tstart, const int* tend, R* scratch, R* grid) {
    const size_t shm = tiled_lds<R, 0>(tg);
    k_spread_tiled<R, 0><<<dim3(tg.ntiles), dim3(256), shm, s>>>(pos, lab, g, tg, nullptr, tstart, tend, scratch,
                                                                  grid);
}
template <class R>
void launch_spread_induced_tiled(hipStream_t s, const R4<R>* pos, const R* mu, PmeGeom<R> g, TileGeom tg,
                                 const int* atoms, const int* tstart, const int* tend, R* scratch, R* grid) {
    const size_t shm = tiled_lds<R, 1>(tg);
    k_spread_tiled<R, 1><<<dim3(tg.ntiles), dim3(256), shm, s>>>(pos, mu, g, tg, atoms, tstart, tend, scratch, grid);
}
template <class R, int MODE> static size_t owner_lds(const TileGeom& tg) {
    constexpr int ND = MODE == 0 ? 3 : 1;
    constexpr int NC = MODE == 0 ? 20 : 3;
    constexpr int K = owner_cand<R, MODE>();
    return (size_t)tg.T[0] * tg.T[1] * tg.T[2] * sizeof(double) + K * (3 * kOrder * (ND + 1) + NC) * sizeof(R) +
           K * 3 * sizeof(int);
}
bool spread_by_owner(const TileGeom& tg) {
    for (int k = 0; k < 3; k++)
        if (tg.T[k] < 8) return false;
    return owner_lds<double, 0>(tg) <= 65536;
}
template <class R>
void launch_spread_fixed_owner(hipStream_t s, const R4<R>* pos, const R* lab, PmeGeom<R> g, TileGeom tg,
                               const int* tstart, const int* tend, R* grid) {
    k_spread_owner<R, 0><<<dim3(tg.ntiles), dim3(256), owner_lds<R, 0>(tg), s>>>(pos, lab, g, tg, nullptr, tstart,
                                                                                   tend, grid);
}
template <class R>
void launch_spread_induced_owner(hipStream_t s, const R4<R>* pos, const R* mu, PmeGeom<R> g, TileGeom tg,
                                 const int* atoms, const int* tstart, const int* tend, R* grid) {
    k_spread_owner<R, 1><<<dim3(tg.ntiles), dim3(256), owner_lds<R, 1>(tg), s>>>(pos, mu, g, tg, atoms, tstart, tend,
                                                                                   grid);
}
template <class R> void launch_fold(hipStream_t s, PmeGeom<R> g, TileGeom tg, const R* scratch, R* grid) {
    size_t total = (size_t)g.n[0] * g.n[1] * g.n[2];
    int blocks = (int)std::min<size_t>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_fold<R>, dim3(blocks), dim3(256), 0, s, g, tg, scratch, grid);
}

bool fold_by_tiles(const TileGeom& tg) {
    for (int k = 0; k < 3; k++)
        if (tg.W[k] != tg.T[k] + 8 || tg.T[k] < 8) return false;
    return tg.tiles_per_axis >= 2;
}
template <class R> void launch_fold_tiles(hipStream_t s, PmeGeom<R> g, TileGeom tg, const R* scratch, R* grid) {
    const FastDiv t2 = make_fastdiv((unsigned)tg.T[2]), t12 = make_fastdiv((unsigned)(tg.T[1] * tg.T[2]));
    hipLaunchKernelGGL(k_fold_tiles<R>, dim3(tg.ntiles), dim3(256), 0, s, g, tg, t2, t12, scratch, grid);
}

template <class R>
void launch_gather_phi(hipStream_t s, int n, const R4<R>* pos, PmeGeom<R> g, const R* grid, R* phi) {
    const size_t threads = 8 * (size_t)n;
    hipLaunchKernelGGL(k_gather_phi<R>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, n, pos, g, grid, phi);
}
template <class R>
void launch_gather_field(hipStream_t s, int natoms, const int* atoms, const R4<R>* pos, PmeGeom<R> g, const R* grid,
                         R* field) {
    if (natoms <= 0) return;
    const size_t threads = 8 * (size_t)natoms;
    hipLaunchKernelGGL(k_gather_field<R>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, natoms, atoms, pos,
                       g, grid, field);
}

#define MPID_INSTANTIATE(R)                                                                                   \
    template void launch_spread_fixed<R>(hipStream_t, int, const R4<R>*, const R*, PmeGeom<R>, R*);           \
    template void launch_spread_induced<R>(hipStream_t, int, const R4<R>*, const R*, PmeGeom<R>, R*);         \
    template void launch_gather_phi<R>(hipStream_t, int, const R4<R>*, PmeGeom<R>, const R*, R*);            \
    template void launch_gather_field<R>(hipStream_t, int, const int*, const R4<R>*, PmeGeom<R>, const R*,   \
                                         R*);                                                                \
    template void launch_spread_fixed_tiled<R>(hipStream_t, const R4<R>*, const R*, PmeGeom<R>, TileGeom,    \
                                               const int*, const int*, R*, R*);                              \
    template void launch_spread_induced_tiled<R>(hipStream_t, const R4<R>*, const R*, PmeGeom<R>, TileGeom,  \
                                                 const int*, const int*, const int*, R*, R*);                \
    template void launch_fold<R>(hipStream_t, PmeGeom<R>, TileGeom, const R*, R*);                          \
    template void launch_spread_fixed_owner<R>(hipStream_t, const R4<R>*, const R*, PmeGeom<R>, TileGeom,    \
                                               const int*, const int*, R*);                                  \
    template void launch_spread_induced_owner<R>(hipStream_t, const R4<R>*, const R*, PmeGeom<R>, TileGeom,  \
                                                 const int*, const int*, const int*, R*);                    \
    template void launch_fold_tiles<R>(hipStream_t, PmeGeom<R>, TileGeom, const R*, R*);
MPID_INSTANTIATE(float)
MPID_INSTANTIATE(double)
#undef MPID_INSTANTIATE

}  // namespace mpid
