// Per-atom kernels of the MPIDForce path on gfx950: spatial sort keys, the
// gather into sorted order, lab-frame rotation of multipoles and
// polarisability tensors, solver vector operations with deterministic FP64
// reductions, reciprocal-space energies/forces/torques, self terms and the
// torque-to-force chain rule.
//
// Follows platforms/reference/src/SimTKReference/MPIDReferenceForce.cpp:
// loadParticleData :282-350, checkChiral :357-397, applyRotationMatrix
// :399-648, recordFixedMultipoleField :4026-4037, the PME self field :2940-2944,
// computeReciprocalSpaceFixedMultipoleForceAndEnergy :3739-3866,
// computeReciprocalSpaceInducedDipoleForceAndEnergy :3871-4024,
// calculatePmeSelfEnergy :4283-4318, calculatePmeSelfTorque :4320-4333,
// mapTorqueToForceForParticle :1895-2110.
#include "engine.h"

#include <cstring>
#include <map>
#include <mutex>
#include <rocprim/rocprim.hpp>

namespace mpid {

typedef V3<double> D3;

__device__ __forceinline__ D3 d3(double4 p) { return mk(p.x, p.y, p.z); }
__device__ __forceinline__ double dnorm(D3& v) {
    double n = sqrt(dot(v, v));
    if (n > 0.0) v = v * (1.0 / n);
    return n;
}

// ---------------------------------------------------------------------------
// sorting
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned spread_bits(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton_keys(int n, const double* __restrict__ pos, Box<double> box, unsigned* keys, int* vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned c[3];
    const D3 fr = to_frac(mk(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]), box);
    const double fs[3] = {fr.x, fr.y, fr.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double f = fs[k];
        f -= floor(f);
        int q = (int)(f * 1024.0);
        c[k] = (unsigned)min(max(q, 0), 1023);
    }
    keys[i] = (spread_bits(c[0]) << 2) | (spread_bits(c[1]) << 1) | spread_bits(c[2]);
    vals[i] = i;
}

__global__ void k_invert_perm(int n, const int* __restrict__ perm, int* inv) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[perm[s]] = s;
}

__global__ void k_gather_sorted(int n, const int* __restrict__ perm, const int* __restrict__ inv,
                                const double* __restrict__ pos, Box<double> box, const double* __restrict__ mol,
                                const int4* __restrict__ axes, const double* __restrict__ thole,
                                const double* __restrict__ pol, const int* __restrict__ partners,
                                const int* __restrict__ npart, int pstride, double4* spos, double* smol,
                                int4* saxes, double* sthole, double* spol, int* spartners, int* snpart) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int a = perm[s];
    // wrapped into the cell (fractional coordinates in [0, 1)); without
    // periodicity (zero box) the positions stay as they are
    D3 w = mk(pos[3 * (size_t)a], pos[3 * (size_t)a + 1], pos[3 * (size_t)a + 2]);
    if (box.L[0] > 0.0) {
        D3 f = to_frac(w, box);
        f.x -= floor(f.x);
        f.y -= floor(f.y);
        f.z -= floor(f.z);
        w = from_frac(f, box);
    }
    const double x[3] = {w.x, w.y, w.z};
    spos[s] = make_double4(x[0], x[1], x[2], mol[(size_t)a * kMP]);
    for (int k = 0; k < kMP; k++) smol[(size_t)s * kMP + k] = mol[(size_t)a * kMP + k];
    int4 ax = axes[a];
    saxes[s] = make_int4(ax.x >= 0 ? inv[ax.x] : -1, ax.y >= 0 ? inv[ax.y] : -1, ax.z >= 0 ? inv[ax.z] : -1, ax.w);
    sthole[s] = thole[a];
    for (int k = 0; k < 3; k++) spol[3 * (size_t)s + k] = pol[3 * (size_t)a + k];
    const int np = npart[a];
    snpart[s] = np;
    for (int k = 0; k < np; k++) {
        int e = partners[(size_t)a * pstride + k];
        spartners[(size_t)s * pstride + k] = (inv[e >> 2] << 2) | (e & 3);
    }
}

// ---------------------------------------------------------------------------
// lab frame
// ---------------------------------------------------------------------------
// Molecular frame axes of atom s (rows x, y, z in lab components); identity
// for NoAxisType.  ZOnly picks its x axis from the lab axes.
__device__ void frame_axes(int s, const double4* __restrict__ pos, int4 ax, const Box<double>& box, D3 e[3]) {
    const int t = ax.w;
    if (t == 5 || ax.x < 0) {
        e[0] = mk(1.0, 0.0, 0.0);
        e[1] = mk(0.0, 1.0, 0.0);
        e[2] = mk(0.0, 0.0, 1.0);
        return;
    }
    const D3 pi = d3(pos[s]);
    D3 vz = min_image(d3(pos[ax.x]) - pi, box);
    dnorm(vz);
    D3 vx;
    if (t == 4) {
        vx = fabs(vz.x) < 0.866 ? mk(1.0, 0.0, 0.0) : mk(0.0, 1.0, 0.0);
    } else {
        vx = min_image(d3(pos[ax.y]) - pi, box);
        if (t == 1) {
            dnorm(vx);
            vz = vz + vx;
            dnorm(vz);
        } else if (t == 2) {
            dnorm(vx);
            D3 vy = min_image(d3(pos[ax.z]) - pi, box);
            dnorm(vy);
            vx = vx + vy;
            dnorm(vx);
        } else if (t == 3) {
            dnorm(vx);
            D3 vy = min_image(d3(pos[ax.z]) - pi, box);
            dnorm(vy);
            vz = vz + vx + vy;
            dnorm(vz);
        }
    }
    vx = vx - vz * dot(vz, vx);
    dnorm(vx);
    e[0] = vx;
    e[1] = cross(vz, vx);
    e[2] = vz;
}

__device__ __forceinline__ int i2(int a, int b) {
    // XX XY XZ YY YZ ZZ
    return a <= b ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : (b == 0 ? a : (b == 1 ? 2 + a : 5));
}
__device__ __forceinline__ int i3(int a, int b, int c) {
    // sort
    int lo = min(a, min(b, c)), hi = max(a, max(b, c)), mid = a + b + c - lo - hi;
    const int t[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}}, {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
                            {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
    return t[lo][mid][hi];
}

// axes: x = z-axis atom, y = x-axis atom, z = y-axis atom, w = axis type (all sorted indices)
__global__ __launch_bounds__(128) void k_rotate(int n, Box<double> box, const double4* __restrict__ pos,
                                                 const int4* __restrict__ axes, const double* __restrict__ mol,
                                                 const double* __restrict__ pol, double* lab, double* labpol,
                                                 double* sqrtpol, double* damp, int* aniso, int* chg) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int4 ax = axes[s];
    double m[kMP];
    bool bare = true;  // charge only: no higher moment, no polarisability
    for (int k = 0; k < kMP; k++) {
        m[k] = mol[(size_t)s * kMP + k];
        if (k > 0 && m[k] != 0.0) bare = false;
    }
    // chirality (ZThenX with a y atom): flip y-odd dipole/quadrupole components
    if (ax.w == 0 && ax.z >= 0) {
        const D3 py = d3(pos[ax.z]);
        D3 ad = min_image(d3(pos[s]) - py, box), bd = min_image(d3(pos[ax.x]) - py, box),
           cd = min_image(d3(pos[ax.y]) - py, box);
        if (dot(cross(bd, cd), ad) < 0.0) {
            m[2] = -m[2];
            m[5] = -m[5];
            m[8] = -m[8];
        }
    }
    D3 e[3];
    frame_axes(s, pos, ax, box, e);
    double E[3][3];  // E[k][a] = component a of molecular axis k
    for (int k = 0; k < 3; k++) {
        E[k][0] = e[k].x;
        E[k][1] = e[k].y;
        E[k][2] = e[k].z;
    }
    double* o = lab + (size_t)s * kMP;
    o[0] = m[0];
    for (int a = 0; a < 3; a++) {
        double v = 0;
        for (int k = 0; k < 3; k++) v += E[k][a] * m[1 + k];
        o[1 + a] = v;
    }
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            double v = 0;
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) v += E[k][a] * E[l][b] * m[4 + i2(k, l)];
            o[4 + i2(a, b)] = v;
        }
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++)
            for (int c = b; c < 3; c++) {
                double v = 0;
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++)
                        for (int q = 0; q < 3; q++) v += E[k][a] * E[l][b] * E[q][c] * m[10 + i3(k, l, q)];
                o[10 + i3(a, b, c)] = v;
            }
    const double p0 = pol[3 * (size_t)s], p1 = pol[3 * (size_t)s + 1], p2 = pol[3 * (size_t)s + 2];
    const double pk[3] = {p0, p1, p2};
    const double sk[3] = {sqrt(fmax(p0, 0.0)), sqrt(fmax(p1, 0.0)), sqrt(fmax(p2, 0.0))};
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            double v = 0, w = 0;
            for (int k = 0; k < 3; k++) {
                v += E[k][a] * E[k][b] * pk[k];
                w += E[k][a] * E[k][b] * sk[k];
            }
            labpol[6 * (size_t)s + i2(a, b)] = v;
            sqrtpol[6 * (size_t)s + i2(a, b)] = w;
        }
    // MPIDForce.h:531 / MPIDReferenceKernels.cpp:134
    damp[s] = pow((p0 + p1 + p2) / 3.0, 1.0 / 6.0);
    aniso[s] = (p0 != p1 || p0 != p2) ? 1 : 0;
    chg[s] = (bare && p0 == 0.0 && p1 == 0.0 && p2 == 0.0) ? 1 : 0;
}

template <class R>
__global__ void k_convert(int n, const double4* __restrict__ pos, const double* __restrict__ lab,
                          const double* __restrict__ thole, const double* __restrict__ damp, R4<R>* rpos, R* rlab,
                          R2<R>* rtd) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double4 p = pos[s];
    rpos[s] = R4<R>{(R)p.x, (R)p.y, (R)p.z, (R)damp[s]};  // w: damping factor
    for (int k = 0; k < kMP; k++) rlab[(size_t)s * kMP + k] = (R)lab[(size_t)s * kMP + k];
    rtd[s] = R2<R>{(R)thole[s], (R)damp[s]};
}

// ---------------------------------------------------------------------------
// fields and solver vectors
// ---------------------------------------------------------------------------
__device__ __forceinline__ D3 sym_mul(const double* S, D3 v) {
    return mk(S[0] * v.x + S[1] * v.y + S[2] * v.z, S[1] * v.x + S[3] * v.y + S[4] * v.z,
              S[2] * v.x + S[4] * v.y + S[5] * v.z);
}

template <class R>
__global__ void k_fixed_field_finish(int n, const R* __restrict__ phi, const double* __restrict__ lab,
                                     const acc_t* __restrict__ acc, FracT<double> A, double selfc, double* E0) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const R* ph = phi + (size_t)s * 35;
    const double* m = lab + (size_t)s * kMP;
    const D3 c1 = cart_vector(mk((double)ph[1], (double)ph[2], (double)ph[3]), A);
    E0[3 * (size_t)s + 0] = -c1.x + selfc * m[1] + acc_get(acc[3 * (size_t)s + 0]);
    E0[3 * (size_t)s + 1] = -c1.y + selfc * m[2] + acc_get(acc[3 * (size_t)s + 1]);
    E0[3 * (size_t)s + 2] = -c1.z + selfc * m[3] + acc_get(acc[3 * (size_t)s + 2]);
}

__global__ void k_solver_init(int n, const double* __restrict__ S, const double* __restrict__ E0, double* b,
                              double* y) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    D3 v = sym_mul(S + 6 * (size_t)s, mk(E0[3 * (size_t)s], E0[3 * (size_t)s + 1], E0[3 * (size_t)s + 2]));
    b[3 * (size_t)s] = y[3 * (size_t)s] = v.x;
    b[3 * (size_t)s + 1] = y[3 * (size_t)s + 1] = v.y;
    b[3 * (size_t)s + 2] = y[3 * (size_t)s + 2] = v.z;
}

template <class R>
__global__ void k_apply_S(int n, const double* __restrict__ S, const double* __restrict__ y, R* mu, double* mud) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    D3 v = sym_mul(S + 6 * (size_t)s, mk(y[3 * (size_t)s], y[3 * (size_t)s + 1], y[3 * (size_t)s + 2]));
    mu[3 * (size_t)s] = (R)v.x;
    mu[3 * (size_t)s + 1] = (R)v.y;
    mu[3 * (size_t)s + 2] = (R)v.z;
    if (mud) {
        mud[3 * (size_t)s] = v.x;
        mud[3 * (size_t)s + 1] = v.y;
        mud[3 * (size_t)s + 2] = v.z;
    }
}

template <class R>
__global__ void k_matvec_finish(int n, const double* __restrict__ S, const acc_t* __restrict__ acc,
                                const R* __restrict__ erec, const R* __restrict__ mu, double selfc,
                                const double* __restrict__ p, double* Ap) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    D3 E;
    E.x = acc_get(acc[3 * (size_t)s]) + (double)erec[3 * (size_t)s] + selfc * (double)mu[3 * (size_t)s];
    E.y = acc_get(acc[3 * (size_t)s + 1]) + (double)erec[3 * (size_t)s + 1] + selfc * (double)mu[3 * (size_t)s + 1];
    E.z = acc_get(acc[3 * (size_t)s + 2]) + (double)erec[3 * (size_t)s + 2] + selfc * (double)mu[3 * (size_t)s + 2];
    D3 v = sym_mul(S + 6 * (size_t)s, E);
    Ap[3 * (size_t)s] = p[3 * (size_t)s] - v.x;
    Ap[3 * (size_t)s + 1] = p[3 * (size_t)s + 1] - v.y;
    Ap[3 * (size_t)s + 2] = p[3 * (size_t)s + 2] - v.z;
}

// Deterministic block reductions: 256 threads, fixed tree order.
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 512;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kRedThreads) void k_dot_partial(int m, const double* __restrict__ a,
                                                             const double* __restrict__ b, double* partial) {
    __shared__ double sh[kRedThreads];
    double v = 0;
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < m; i += kRedBlocks * kRedThreads) v += a[i] * b[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// Sums kRedBlocks partials per output slot; nout outputs stored back to back.
__global__ __launch_bounds__(kRedThreads) void k_reduce_partials(const double* __restrict__ partial, int nout,
                                                                 double* result, const double* skip) {
    if (skipped(skip)) return;
    __shared__ double sh[kRedThreads];
    for (int o = 0; o < nout; o++) {
        double v = 0;
        for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) v += partial[o * kRedBlocks + i];
        v = block_sum(v, sh);
        if (threadIdx.x == 0) result[o] = v;
    }
}

// x += a p, r -= a Ap with a = rr / pAp; partial sums of r.r and |S r|^2
__global__ __launch_bounds__(kRedThreads) void k_cg_step(int n, const double* __restrict__ S,
                                                         const double* __restrict__ scal,
                                                         const double* __restrict__ p, const double* __restrict__ Ap,
                                                         double* x, double* r, double* partial) {
    if (scal[kCgStop] != 0.0) return;
    __shared__ double sh[kRedThreads];
    const double pAp = scal[1];
    const double a = pAp != 0.0 ? scal[0] / pAp : 0.0;
    double rr = 0, ss = 0;
    for (int s = blockIdx.x * kRedThreads + threadIdx.x; s < n; s += kRedBlocks * kRedThreads) {
        D3 rv;
        for (int k = 0; k < 3; k++) {
            const size_t id = 3 * (size_t)s + k;
            x[id] += a * p[id];
            const double nr = r[id] - a * Ap[id];
            r[id] = nr;
            rr += nr * nr;
        }
        rv = mk(r[3 * (size_t)s], r[3 * (size_t)s + 1], r[3 * (size_t)s + 2]);
        D3 sr = sym_mul(S + 6 * (size_t)s, rv);
        ss += dot(sr, sr);
    }
    rr = block_sum(rr, sh);
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = rr;
        partial[kRedBlocks + blockIdx.x] = ss;
    }
}

__global__ __launch_bounds__(kRedThreads) void k_residual_norm(int n, const double* __restrict__ S,
                                                               const double* __restrict__ r, double* partial) {
    __shared__ double sh[kRedThreads];
    double rr = 0, ss = 0;
    for (int s = blockIdx.x * kRedThreads + threadIdx.x; s < n; s += kRedBlocks * kRedThreads) {
        D3 rv = mk(r[3 * (size_t)s], r[3 * (size_t)s + 1], r[3 * (size_t)s + 2]);
        rr += dot(rv, rv);
        D3 sr = sym_mul(S + 6 * (size_t)s, rv);
        ss += dot(sr, sr);
    }
    rr = block_sum(rr, sh);
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = rr;
        partial[kRedBlocks + blockIdx.x] = ss;
    }
}

// The solver's operator output fused with the CG inner product: Ap as in
// k_matvec_finish and the partial sums of p.Ap per block (kRedBlocks blocks).
template <class R>
__global__ __launch_bounds__(kRedThreads) void k_matvec_finish_dot(int n, const double* __restrict__ S,
                                                                   const acc_t* __restrict__ acc,
                                                                   const R* __restrict__ erec,
                                                                   const R* __restrict__ mu, double selfc,
                                                                   const double* __restrict__ p, double* Ap,
                                                                   double* partial, const double* __restrict__ scal) {
    if (scal[kCgStop] != 0.0) return;
    __shared__ double sh[kRedThreads];
    double v = 0;
    for (int s = blockIdx.x * kRedThreads + threadIdx.x; s < n; s += kRedBlocks * kRedThreads) {
        D3 E;
        E.x = acc_get(acc[3 * (size_t)s]) + (double)erec[3 * (size_t)s] + selfc * (double)mu[3 * (size_t)s];
        E.y = acc_get(acc[3 * (size_t)s + 1]) + (double)erec[3 * (size_t)s + 1] +
              selfc * (double)mu[3 * (size_t)s + 1];
        E.z = acc_get(acc[3 * (size_t)s + 2]) + (double)erec[3 * (size_t)s + 2] +
              selfc * (double)mu[3 * (size_t)s + 2];
        const D3 w = sym_mul(S + 6 * (size_t)s, E);
        const D3 pv = mk(p[3 * (size_t)s], p[3 * (size_t)s + 1], p[3 * (size_t)s + 2]);
        const D3 a = mk(pv.x - w.x, pv.y - w.y, pv.z - w.z);
        Ap[3 * (size_t)s] = a.x;
        Ap[3 * (size_t)s + 1] = a.y;
        Ap[3 * (size_t)s + 2] = a.z;
        v += dot(pv, a);
    }
    v = block_sum(v, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// The solver's stopping test, the same expression as the host's report:
// converged when the RMS of S r, in Debye, is below stop_eps.
__device__ __forceinline__ double cg_stopped(double ss, int n, double stop_eps) {
    return kDebye * sqrt(ss / (double)n) >= stop_eps ? 0.0 : 1.0;
}

// Publishes check c of a solve (the state after c iterations): scal[kCgCount]
// = c, scal[kCgStop] = the stopping test, and in host slot c & 1 the pair
// (|S r|^2, stop).  Slots alternate so that the host can read check c while
// the iteration after it is already running.
__device__ __forceinline__ void cg_publish(double* scal, double* hcheck, double c, double ss, double stop) {
    scal[kCgCount] = c;
    scal[kCgStop] = stop;
    double* slot = hcheck + 2 * ((int)c & 1);
    slot[0] = ss;
    slot[1] = stop;
}

// Check 0 of a solve, from the initial residual's |S r|^2 in scal[3].
__global__ void k_cg_begin(double* scal, double* hcheck, int n, double stop_eps) {
    if (threadIdx.x != 0) return;
    const double ss = scal[3];
    cg_publish(scal, hcheck, 0.0, ss, cg_stopped(ss, n, stop_eps));
}

// One block: the partial sums of k_cg_step give rr_new (scal[2], scal[0]) and
// ss (scal[3]); beta = rr_new / rr_old goes to scal[7], then the iteration's
// check is published.  Thread 0 reads the old scal[0] before it writes the new
// one.
__global__ __launch_bounds__(kRedThreads) void k_cg_reduce_beta(const double* __restrict__ partial, double* scal,
                                                                double* hcheck, int n, double stop_eps) {
    if (scal[kCgStop] != 0.0) return;
    __shared__ double sh[kRedThreads];
    double rr = 0, ss = 0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) {
        rr += partial[i];
        ss += partial[kRedBlocks + i];
    }
    rr = block_sum(rr, sh);
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) {
        const double old = scal[0];
        scal[7] = old != 0.0 ? rr / old : 0.0;
        scal[0] = rr;
        scal[2] = rr;
        scal[3] = ss;
        cg_publish(scal, hcheck, scal[kCgCount] + 1.0, ss, cg_stopped(ss, n, stop_eps));
    }
}

// p = r + beta p (beta = scal[7]), fused with the next operator input mu = S p.
// Skipped once the solve has stopped, also after the iteration that stops it:
// the solver does not read p or mu after its last check.
template <class R>
__global__ void k_cg_direction_apply(int n, const double* __restrict__ S, const double* __restrict__ scal,
                                     const double* __restrict__ r, double* p, R* mu) {
    if (scal[kCgStop] != 0.0) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double beta = scal[7];
    D3 pv;
    pv.x = r[3 * (size_t)s] + beta * p[3 * (size_t)s];
    pv.y = r[3 * (size_t)s + 1] + beta * p[3 * (size_t)s + 1];
    pv.z = r[3 * (size_t)s + 2] + beta * p[3 * (size_t)s + 2];
    p[3 * (size_t)s] = pv.x;
    p[3 * (size_t)s + 1] = pv.y;
    p[3 * (size_t)s + 2] = pv.z;
    const D3 v = sym_mul(S + 6 * (size_t)s, pv);
    mu[3 * (size_t)s] = (R)v.x;
    mu[3 * (size_t)s + 1] = (R)v.y;
    mu[3 * (size_t)s + 2] = (R)v.z;
}

template <class R>
__device__ void cart_potential(const R* ph, const FracT<double>& A, double* c1, double* c2, double* c3);

// ---- Extrapolated (OPT) polarization --------------------------------------
// convergeInduceDipolesByExtrapolation, MPIDReferenceForce.cpp:1125-1180.
// One order: the field of mu_(m) at every atom, E = real space (acc) +
// reciprocal (-A phi_1) + selfc mu_(m), and its gradient G = real space
// (gacc) - A phi_2 A^T (A the grid-coordinate transform; :4094-4127), kept
// for the response forces; then mu_(m+1) = alpha_lab E, in FP64 and in the
// precision of the next order's kernels.
template <class R>
__global__ void k_extrap_order(int n, const acc_t* __restrict__ field, const acc_t* __restrict__ gacc,
                               const R* __restrict__ phidp, FracT<double> A, double selfc,
                               const double* __restrict__ labpol, const double* __restrict__ mu, double* E,
                               double* G, double* mu_next, R* mu_next_r) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const R* ph = phidp + 35 * (size_t)s;
    const size_t i3 = 3 * (size_t)s, i6 = 6 * (size_t)s;
    double c1[3], c2[6], c3[10];
    cart_potential<R>(ph, A, c1, c2, c3);
    D3 e;
    e.x = acc_get(field[i3]) - c1[0] + selfc * mu[i3];
    e.y = acc_get(field[i3 + 1]) - c1[1] + selfc * mu[i3 + 1];
    e.z = acc_get(field[i3 + 2]) - c1[2] + selfc * mu[i3 + 2];
    E[i3] = e.x;
    E[i3 + 1] = e.y;
    E[i3 + 2] = e.z;
    // gradient order xx yy zz xy xz yz; c2 is XX XY XZ YY YZ ZZ
    G[i6 + 0] = acc_get(gacc[i6 + 0]) - c2[0];
    G[i6 + 1] = acc_get(gacc[i6 + 1]) - c2[3];
    G[i6 + 2] = acc_get(gacc[i6 + 2]) - c2[5];
    G[i6 + 3] = acc_get(gacc[i6 + 3]) - c2[1];
    G[i6 + 4] = acc_get(gacc[i6 + 4]) - c2[2];
    G[i6 + 5] = acc_get(gacc[i6 + 5]) - c2[4];
    const D3 m = sym_mul(labpol + i6, e);
    mu_next[i3] = m.x;
    mu_next[i3 + 1] = m.y;
    mu_next[i3 + 2] = m.z;
    mu_next_r[i3] = (R)m.x;
    mu_next_r[i3 + 1] = (R)m.y;
    mu_next_r[i3 + 2] = (R)m.z;
}

// mu = sum_o c'_o mu_(o), c'_o = sum_{j >= o} c_j (setExtrapolationCoefficients,
// MPIDReferenceForce.cpp:173-183; the sum at :1170-1177).
template <class R>
__global__ void k_extrap_combine(int n, ExtrapCoefficients cp, const double* __restrict__ mus, double* mud, R* mu_r) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n3 = 3 * (size_t)n;
    if (id >= (int)n3) return;
    double v = 0.0;
    for (int o = 0; o < cp.n; o++) v += cp.c[o] * mus[o * n3 + id];
    mud[id] = v;
    mu_r[id] = (R)v;
}

// The dipole-response forces and torques, MPIDReferenceForce.cpp:4955-4985:
// for l + m <= N - 2, c'_(l+m+1) times mu_(l) contracted with the gradient
// of the field of mu_(m) (force) or crossed with that field (torque on
// anisotropic sites).  Terms with |c'| < 1e-6 are skipped as there.
__global__ void k_extrap_response(int n, ExtrapCoefficients cp, const double* __restrict__ mus,
                                  const double* __restrict__ Es, const double* __restrict__ Gs,
                                  const int* __restrict__ aniso, acc_t* force, acc_t* torque) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t n3 = 3 * (size_t)n, n6 = 6 * (size_t)n;
    const bool an = aniso[s] != 0;
    D3 f = mk(0.0, 0.0, 0.0), t = mk(0.0, 0.0, 0.0);
    for (int l = 0; l < cp.n - 1; l++) {
        const double* u = mus + l * n3 + 3 * (size_t)s;
        for (int m = 0; m < cp.n - 1 - l; m++) {
            const double p = cp.c[l + m + 1];
            if (fabs(p) < 1e-6) continue;
            const double pf = p * kElectric;
            const double* g = Gs + m * n6 + 6 * (size_t)s;
            f.x += pf * (u[0] * g[0] + u[1] * g[3] + u[2] * g[4]);
            f.y += pf * (u[0] * g[3] + u[1] * g[1] + u[2] * g[5]);
            f.z += pf * (u[0] * g[4] + u[1] * g[5] + u[2] * g[2]);
            if (an) {
                const double* e = Es + m * n3 + 3 * (size_t)s;
                t.x += pf * (u[1] * e[2] - u[2] * e[1]);
                t.y += pf * (u[2] * e[0] - u[0] * e[2]);
                t.z += pf * (u[0] * e[1] - u[1] * e[0]);
            }
        }
    }
    acc_add(&force[3 * (size_t)s + 0], f.x);
    acc_add(&force[3 * (size_t)s + 1], f.y);
    acc_add(&force[3 * (size_t)s + 2], f.z);
    if (an) {
        acc_add(&torque[3 * (size_t)s + 0], t.x);
        acc_add(&torque[3 * (size_t)s + 1], t.y);
        acc_add(&torque[3 * (size_t)s + 2], t.z);
    }
}

// ---- getters after an evaluation -----------------------------------------
// getElectrostaticPotential, MPIDReferenceForce.cpp:2464-2531: at each point,
// the Coulomb constant times the bare potential of every particle's charge,
// dipole plus induced dipole and quadrupole (no octopoles), with the
// separation particle minus point taken to its nearest image.  One thread per
// point; the particles pass through LDS in groups of kPotThreads.
constexpr int kPotThreads = 256;
__global__ __launch_bounds__(kPotThreads) void k_potential_points(int n, int npts, const double4* __restrict__ pos,
                                                                  const double* __restrict__ lab,
                                                                  const double* __restrict__ mu, Box<double> box,
                                                                  const double* __restrict__ points, double* out) {
    __shared__ double sh[13][kPotThreads];  // x y z, q, d + mu (3), Q xx xy xz yy yz zz
    const int g = blockIdx.x * kPotThreads + threadIdx.x;
    const bool gv = g < npts;
    const V3<double> pt = gv ? mk(points[3 * (size_t)g], points[3 * (size_t)g + 1], points[3 * (size_t)g + 2])
                             : mk(0.0, 0.0, 0.0);
    double phi = 0.0;
    for (int base = 0; base < n; base += kPotThreads) {
        const int a = base + threadIdx.x;
        if (a < n) {
            const double4 p = pos[a];
            const double* m = lab + (size_t)a * kMP;
            sh[0][threadIdx.x] = p.x;
            sh[1][threadIdx.x] = p.y;
            sh[2][threadIdx.x] = p.z;
            sh[3][threadIdx.x] = m[0];
            for (int k = 0; k < 3; k++) sh[4 + k][threadIdx.x] = m[1 + k] + mu[3 * (size_t)a + k];
            for (int k = 0; k < 6; k++) sh[7 + k][threadIdx.x] = m[4 + k];
        }
        __syncthreads();
        const int cnt = min(kPotThreads, n - base);
        for (int j = 0; j < cnt; j++) {
            const V3<double> d =
                min_image(mk(sh[0][j] - pt.x, sh[1][j] - pt.y, sh[2][j] - pt.z), box);
            const double r2 = dot(d, d);
            const double rr1 = 1.0 / sqrt(r2), rr3 = rr1 * rr1 * rr1, rr5 = 3.0 * rr3 * rr1 * rr1;
            const double qx = sh[7][j] * d.x + sh[8][j] * d.y + sh[9][j] * d.z;
            const double qy = sh[8][j] * d.x + sh[10][j] * d.y + sh[11][j] * d.z;
            const double qz = sh[9][j] * d.x + sh[11][j] * d.y + sh[12][j] * d.z;
            phi += sh[3][j] * rr1 - (sh[4][j] * d.x + sh[5][j] * d.y + sh[6][j] * d.z) * rr3 +
                   (d.x * qx + d.y * qy + d.z * qz) * rr5;
        }
        __syncthreads();
    }
    if (gv) out[g] = kElectric * phi;
}

// getSystemMultipoleMoments, MPIDReferenceForce.cpp:2352-2462, one block in
// a fixed order: the centre of mass of the given masses, then about it the
// net charge, the dipole (with the induced dipoles) and the traced second
// moment (xx xy xz yy yz zz), and the sum of the atomic quadrupoles.  Raw
// positions and masses are in the caller's order (perm: sorted -> caller).
// out: [0] mass, [1..3] mass-weighted position, [4] charge, [5..7] dipole,
// [8..13] second moment, [14..19] quadrupole sum.
constexpr int kMomThreads = 1024;
__device__ __forceinline__ double block_sum_1024(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = kMomThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
__global__ __launch_bounds__(kMomThreads) void k_multipole_moments(int n, const int* __restrict__ perm,
                                                                   const double* __restrict__ pos,
                                                                   const double* __restrict__ masses,
                                                                   const double* __restrict__ lab,
                                                                   const double* __restrict__ mu, double* out) {
    __shared__ double sh[kMomThreads];
    double m = 0, mx = 0, my = 0, mz = 0;
    for (int s = threadIdx.x; s < n; s += kMomThreads) {
        const int a = perm[s];
        const double w = masses[a];
        m += w;
        mx += w * pos[3 * (size_t)a];
        my += w * pos[3 * (size_t)a + 1];
        mz += w * pos[3 * (size_t)a + 2];
    }
    m = block_sum_1024(m, sh);
    mx = block_sum_1024(mx, sh);
    my = block_sum_1024(my, sh);
    mz = block_sum_1024(mz, sh);
    const double c[3] = {m > 0.0 ? mx / m : 0.0, m > 0.0 ? my / m : 0.0, m > 0.0 ? mz / m : 0.0};
    double acc[16];
    for (int k = 0; k < 16; k++) acc[k] = 0.0;
    for (int s = threadIdx.x; s < n; s += kMomThreads) {
        const int a = perm[s];
        const double* l = lab + (size_t)s * kMP;
        const double x[3] = {pos[3 * (size_t)a] - c[0], pos[3 * (size_t)a + 1] - c[1], pos[3 * (size_t)a + 2] - c[2]};
        const double d[3] = {l[1] + mu[3 * (size_t)s], l[2] + mu[3 * (size_t)s + 1], l[3] + mu[3 * (size_t)s + 2]};
        const double q = l[0];
        acc[0] += q;
        for (int k = 0; k < 3; k++) acc[1 + k] += x[k] * q + d[k];
        const int ia[6] = {0, 0, 0, 1, 1, 2}, ic[6] = {0, 1, 2, 1, 2, 2};
        for (int k = 0; k < 6; k++)
            acc[4 + k] += x[ia[k]] * x[ic[k]] * q + x[ia[k]] * d[ic[k]] + x[ic[k]] * d[ia[k]];
        for (int k = 0; k < 6; k++) acc[10 + k] += l[4 + k];
    }
    for (int k = 0; k < 16; k++) acc[k] = block_sum_1024(acc[k], sh);
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = mx;
        out[2] = my;
        out[3] = mz;
        for (int k = 0; k < 16; k++) out[4 + k] = acc[k];
    }
}

__global__ void k_axpy(int m, double a, const double* __restrict__ x, double* y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) y[i] += a * x[i];
}
__global__ void k_sub(int m, const double* __restrict__ a, const double* __restrict__ b, double* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) out[i] = a[i] - b[i];
}

__global__ void k_zero_acc(acc_t* p, size_t count, const double* skip) {
    if (skipped(skip)) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}
template <class R> __global__ void k_fill(R* p, size_t count, R v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// reciprocal-space energies, forces, torques and self terms (per atom)
// ---------------------------------------------------------------------------
// phi index of the derivative with exponents (a, b, c), a+b+c <= 4
__device__ __forceinline__ int phi_idx(int a, int b, int c) {
    const int t[5][5][5] = {
        {{0, 3, 6, 12, 22}, {2, 9, 18, 28, -1}, {5, 16, 31, -1, -1}, {11, 26, -1, -1, -1}, {21, -1, -1, -1, -1}},
        {{1, 8, 17, 27, -1}, {7, 19, 34, -1, -1}, {15, 33, -1, -1, -1}, {25, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}},
        {{4, 14, 30, -1, -1}, {13, 32, -1, -1, -1}, {29, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}},
        {{10, 24, -1, -1, -1}, {23, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}},
        {{20, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}}};
    return t[a][b][c];
}

// exponents (x, y, z counts) of multipole component k of a 20-record
__device__ __forceinline__ void comp_exp(int k, int e[3]) {
    const int t[20][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1},
                          {0, 2, 0}, {0, 1, 1}, {0, 0, 2}, {3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0},
                          {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}};
    e[0] = t[k][0];
    e[1] = t[k][1];
    e[2] = t[k][2];
}
// multiplicity of the unique component among all index orderings
__device__ __forceinline__ double comp_mult(int k) {
    const double t[20] = {1, 1, 1, 1, 1, 2, 2, 1, 2, 1, 1, 3, 3, 3, 6, 3, 1, 3, 3, 1};
    return t[k];
}
__device__ __forceinline__ int levi(int a, int b, int c) {
    if (a == b || b == c || a == c) return 0;
    return ((a == 0 && b == 1) || (a == 1 && b == 2) || (a == 2 && b == 0)) ? 1 : -1;
}

// Torque -dE/dtheta of moments (d, Q, O) in a potential whose Cartesian
// derivative tensors are c1, c2 (6), c3 (10); E = d.c1 + Q:c2 + O:c3 over all
// index orderings.
__device__ D3 multipole_torque(const double* d, const double* Q, const double* O, const double* c1, const double* c2,
                               const double* c3) {
    double t[3];
    for (int k = 0; k < 3; k++) {
        double s = 0;
        for (int a = 0; a < 3; a++)
            for (int i = 0; i < 3; i++) s += levi(a, k, i) * d[i] * c1[a];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double dq = 0;
                for (int i = 0; i < 3; i++) dq += levi(a, k, i) * Q[i2(i, b)] + levi(b, k, i) * Q[i2(a, i)];
                s += dq * c2[i2(a, b)];
            }
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    double dq = 0;
                    for (int i = 0; i < 3; i++)
                        dq += levi(a, k, i) * O[i3(i, b, c)] + levi(b, k, i) * O[i3(a, i, c)] +
                              levi(c, k, i) * O[i3(a, b, i)];
                    s += dq * c3[i3(a, b, c)];
                }
        t[k] = -s;
    }
    return mk(t[0], t[1], t[2]);
}

// Cartesian derivative tensors of the potential from grid-coordinate ones:
// c1_j = sum_k A_jk phi_k, c2_ij = sum_kl A_ik A_jl phi_kl, c3_ijn =
// sum_klm A_ik A_jl A_nm phi_klm (transformPotentialToCartesianCoordinates,
// MPIDReferenceForce.cpp:3171-3267); c2 in the order XX XY XZ YY YZ ZZ, c3 in
// the octopole order.  A rectangular box has A diagonal.
template <class R>
__device__ void cart_potential(const R* ph, const FracT<double>& A, double* c1, double* c2, double* c3) {
    const double a[3][3] = {{A.a00, 0.0, 0.0}, {A.a10, A.a11, 0.0}, {A.a20, A.a21, A.a22}};
    if (!A.tri) {
        const double g[3] = {A.a00, A.a11, A.a22};
        for (int i = 0; i < 3; i++) {
            int e[3] = {0, 0, 0};
            e[i]++;
            c1[i] = g[i] * (double)ph[phi_idx(e[0], e[1], e[2])];
        }
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                int e[3] = {0, 0, 0};
                e[i]++;
                e[j]++;
                c2[i2(i, j)] = g[i] * g[j] * (double)ph[phi_idx(e[0], e[1], e[2])];
            }
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++)
                for (int k = j; k < 3; k++) {
                    int e[3] = {0, 0, 0};
                    e[i]++;
                    e[j]++;
                    e[k]++;
                    c3[i3(i, j, k)] = g[i] * g[j] * g[k] * (double)ph[phi_idx(e[0], e[1], e[2])];
                }
        return;
    }
    double p1[3], p2[3][3], p3[3][3][3];
    for (int k = 0; k < 3; k++) {
        int e[3] = {0, 0, 0};
        e[k]++;
        p1[k] = (double)ph[phi_idx(e[0], e[1], e[2])];
        for (int l = 0; l < 3; l++) {
            int f[3] = {e[0], e[1], e[2]};
            f[l]++;
            p2[k][l] = (double)ph[phi_idx(f[0], f[1], f[2])];
            for (int m = 0; m < 3; m++) {
                int h[3] = {f[0], f[1], f[2]};
                h[m]++;
                p3[k][l][m] = (double)ph[phi_idx(h[0], h[1], h[2])];
            }
        }
    }
    for (int i = 0; i < 3; i++) {
        double v = 0;
        for (int k = 0; k < 3; k++) v += a[i][k] * p1[k];
        c1[i] = v;
    }
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) {
            double v = 0;
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) v += a[i][k] * a[j][l] * p2[k][l];
            c2[i2(i, j)] = v;
        }
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++)
            for (int n = j; n < 3; n++) {
                double v = 0;
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++)
                        for (int m = 0; m < 3; m++) v += a[i][k] * a[j][l] * a[n][m] * p3[k][l][m];
                c3[i3(i, j, n)] = v;
            }
}

template <class R>
__global__ __launch_bounds__(128) void k_recip_atoms(int n, int mutual, const double* __restrict__ phi,
                                                      const R* __restrict__ phidp, const double* __restrict__ lab,
                                                      const double* __restrict__ mu, const int* __restrict__ aniso,
                                                      FracT<double> A, double alpha, acc_t* force, acc_t* torque,
                                                      acc_t* energy) {
    __shared__ double sh[128];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double en = 0;
    if (s < n) {
        const double* m = lab + (size_t)s * kMP;
        const double* ph = phi + (size_t)s * 35;
        const R* pd = phidp + (size_t)s * 35;
        const double u[3] = {mu[3 * (size_t)s], mu[3 * (size_t)s + 1], mu[3 * (size_t)s + 2]};
        const bool an = aniso[s] != 0;
        // fractional multipoles
        double mf[20];
        frac_moments<double>(m, A, mf);
        // fixed: energy 1/2 eps M.phi, force -eps grad, torque
        double ef = 0, f[3] = {0, 0, 0}, fi[3] = {0, 0, 0};
        for (int k = 0; k < 20; k++) {
            int e[3];
            comp_exp(k, e);
            ef += mf[k] * (double)ph[phi_idx(e[0], e[1], e[2])];
            for (int a = 0; a < 3; a++) {
                int e2[3] = {e[0], e[1], e[2]};
                e2[a]++;
                const int id = phi_idx(e2[0], e2[1], e2[2]);
                f[a] += mf[k] * (double)ph[id];
                fi[a] += 2.0 * mf[k] * (double)pd[id];
            }
        }
        // induced: energy 1/4 eps 2 u'.grad phi, force
        const D3 ufv = frac_vector(mk(u[0], u[1], u[2]), A);
        const double uf[3] = {ufv.x, ufv.y, ufv.z};
        double ei = 0;
        for (int k = 0; k < 3; k++) {
            int e[3] = {0, 0, 0};
            e[k]++;
            ei += 2.0 * uf[k] * (double)ph[phi_idx(e[0], e[1], e[2])];
            for (int a = 0; a < 3; a++) {
                int e2[3] = {0, 0, 0};
                e2[k]++;
                e2[a]++;
                const int id = phi_idx(e2[0], e2[1], e2[2]);
                fi[a] += 2.0 * uf[k] * (double)ph[id];
                if (mutual) fi[a] += 2.0 * uf[k] * (double)pd[id];
            }
        }
        const D3 Fc = cart_vector(mk(kElectric * f[0] + 0.5 * kElectric * fi[0], kElectric * f[1] + 0.5 * kElectric * fi[1],
                                     kElectric * f[2] + 0.5 * kElectric * fi[2]),
                                  A);
        const double F[3] = {-Fc.x, -Fc.y, -Fc.z};
        // torques
        double c1[3], c2[6], c3[10];
        double dd[3] = {m[1], m[2], m[3]};
        if (an) {
            dd[0] += u[0];
            dd[1] += u[1];
            dd[2] += u[2];
        }
        cart_potential<double>(ph, A, c1, c2, c3);
        D3 T = multipole_torque(dd, m + 4, m + 10, c1, c2, c3) * kElectric;
        double di[3] = {m[1], m[2], m[3]};
        if (mutual && an) {
            di[0] += u[0];
            di[1] += u[1];
            di[2] += u[2];
        }
        cart_potential<R>(pd, A, c1, c2, c3);
        T = T + multipole_torque(di, m + 4, m + 10, c1, c2, c3) * kElectric;
        // self energy and torque
        const double sqp = 1.7320508075688772;
        const double sq[5] = {3.0 * m[9], 2.0 * sqp * m[6], 2.0 * sqp * m[8], sqp * (m[4] - m[7]), 2.0 * sqp * m[5]};
        const double r32 = 1.2247448713915890, r35 = 0.7745966692414834, r110 = 0.31622776601683794;
        const double so[7] = {15.0 * m[19], 15.0 * r32 * m[15], 15.0 * r32 * m[18], 15.0 * r35 * (m[12] - m[17]),
                              30.0 * r35 * m[14], 15.0 * r110 * (m[10] - 3.0 * m[13]),
                              15.0 * r110 * (3.0 * m[11] - m[16])};
        double qq = 0, oo = 0;
        for (int k = 0; k < 5; k++) qq += sq[k] * sq[k];
        for (int k = 0; k < 7; k++) oo += so[k] * so[k];
        const double a2 = alpha * alpha;
        const double dud = m[1] * (m[1] + u[0]) + m[2] * (m[2] + u[1]) + m[3] * (m[3] + u[2]);
        const double eself = -alpha * kElectric / kSqrtPi *
                             (m[0] * m[0] + (2.0 / 3.0) * a2 * dud + (4.0 / 15.0) * a2 * a2 * qq +
                              (8.0 / 105.0) * a2 * a2 * a2 * oo);
        if (!an) {
            const double term = (2.0 / 3.0) * kElectric * alpha * a2 / kSqrtPi;
            T = T + cross(mk(m[1], m[2], m[3]), mk(2.0 * u[0], 2.0 * u[1], 2.0 * u[2])) * term;
        }
        en = 0.5 * kElectric * ef + 0.25 * kElectric * ei + eself;
        acc_add(&force[3 * (size_t)s + 0], F[0]);
        acc_add(&force[3 * (size_t)s + 1], F[1]);
        acc_add(&force[3 * (size_t)s + 2], F[2]);
        acc_add(&torque[3 * (size_t)s + 0], T.x);
        acc_add(&torque[3 * (size_t)s + 1], T.y);
        acc_add(&torque[3 * (size_t)s + 2], T.z);
    }
    sh[threadIdx.x] = en;
    __syncthreads();
    for (int off = 64; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc_add(energy + (blockIdx.x & (kEnergySlots - 1)), sh[0]);
}

// ---------------------------------------------------------------------------
// energy gradients (per atom)
// ---------------------------------------------------------------------------
// Reciprocal and self parts of the potential vector dE/dm of each atom (the
// 20 stored lab components, induced dipoles fixed), added to pot.  The
// permanent reciprocal energy is 1/2 eps sum mf.phi over the fractional
// moments mf_k = c_k m_k, and the induced dipoles' reciprocal interaction
// with the moments eps sum mf.phidp (phidp: the potential of the induced
// dipoles), so dE/dm_k = eps c_k (phi + phidp) at the index of component k.
// Self terms: the derivatives of calculatePmeSelfEnergy's expression
// (MPIDReferenceForce.cpp:4283-4318, as in k_recip_atoms) with the dipole-
// induced dipole product counted twice, as -mu.E_self carries it at fixed
// induced dipoles.
template <class R>
__global__ __launch_bounds__(128) void k_recip_potential(int n, int pme, const double* __restrict__ phi,
                                                          const R* __restrict__ phidp, const double* __restrict__ lab,
                                                          const double* __restrict__ mu, FracT<double> A,
                                                          double alpha, acc_t* pot) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double* m = lab + (size_t)s * kMP;
    double v[kMP];
    for (int k = 0; k < kMP; k++) v[k] = 0.0;
    if (pme) {
        // the Cartesian derivatives of phi + phidp, times the number of index
        // orderings each stored component stands for
        const double* ph = phi + (size_t)s * 35;
        const R* pd = phidp + (size_t)s * 35;
        double c1[3], c2[6], c3[10], d1[3], d2[6], d3[10];
        cart_potential<double>(ph, A, c1, c2, c3);
        cart_potential<R>(pd, A, d1, d2, d3);
        v[0] = kElectric * (ph[0] + (double)pd[0]);
        for (int k = 0; k < 3; k++) v[1 + k] = kElectric * (c1[k] + d1[k]);
        for (int k = 0; k < 6; k++) v[4 + k] = kElectric * comp_mult(4 + k) * (c2[k] + d2[k]);
        for (int k = 0; k < 10; k++) v[10 + k] = kElectric * comp_mult(10 + k) * (c3[k] + d3[k]);
    }
    if (alpha > 0.0) {
        const double c = -alpha * kElectric / kSqrtPi, a2 = alpha * alpha;
        v[0] += 2.0 * c * m[0];
        for (int a = 0; a < 3; a++) v[1 + a] += c * (4.0 / 3.0) * a2 * (m[1 + a] + mu[3 * (size_t)s + a]);
        // the spherical quadrupole and octopole norms of the self energy
        const double s3 = 1.7320508075688772;
        const double sq[5] = {3.0 * m[9], 2.0 * s3 * m[6], 2.0 * s3 * m[8], s3 * (m[4] - m[7]), 2.0 * s3 * m[5]};
        const double r32 = 1.2247448713915890, r35 = 0.7745966692414834, r110 = 0.31622776601683794;
        const double so[7] = {15.0 * m[19], 15.0 * r32 * m[15], 15.0 * r32 * m[18], 15.0 * r35 * (m[12] - m[17]),
                              30.0 * r35 * m[14], 15.0 * r110 * (m[10] - 3.0 * m[13]),
                              15.0 * r110 * (3.0 * m[11] - m[16])};
        const double cq = c * (4.0 / 15.0) * a2 * a2 * 2.0, co = c * (8.0 / 105.0) * a2 * a2 * a2 * 2.0;
        v[9] += cq * 3.0 * sq[0];
        v[6] += cq * 2.0 * s3 * sq[1];
        v[8] += cq * 2.0 * s3 * sq[2];
        v[4] += cq * s3 * sq[3];
        v[7] -= cq * s3 * sq[3];
        v[5] += cq * 2.0 * s3 * sq[4];
        v[19] += co * 15.0 * so[0];
        v[15] += co * 15.0 * r32 * so[1];
        v[18] += co * 15.0 * r32 * so[2];
        v[12] += co * 15.0 * r35 * so[3];
        v[17] -= co * 15.0 * r35 * so[3];
        v[14] += co * 30.0 * r35 * so[4];
        v[10] += co * 15.0 * r110 * so[5];
        v[13] -= co * 45.0 * r110 * so[5];
        v[11] += co * 45.0 * r110 * so[6];
        v[16] -= co * 15.0 * r110 * so[6];
    }
    for (int k = 0; k < kMP; k++)
        if (v[k] != 0.0) acc_add(&pot[kMP * (size_t)s + k], v[k]);
}

// The total field at each atom, E0 + T mu (mutual polarization: the real-space
// pairs in acc, the reciprocal field -grad phidp and the self term), or E0
// alone (direct polarization, where mu = alpha E0).
template <class R>
__global__ void k_total_field(int n, int mutual, const double* __restrict__ E0, const acc_t* __restrict__ acc,
                              const R* __restrict__ phidp, FracT<double> A, const double* __restrict__ mu,
                              double selfc, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const R* pd = phidp + (size_t)s * 35;
    const D3 gp = cart_vector(mk((double)pd[1], (double)pd[2], (double)pd[3]), A);
    const double grad[3] = {gp.x, gp.y, gp.z};
    for (int k = 0; k < 3; k++) {
        const size_t id = 3 * (size_t)s + k;
        double e = E0[id];
        if (mutual) e += acc_get(acc[id]) - grad[k] + selfc * mu[id];
        out[id] = e;
    }
}

// Per-atom gradients of the energy with respect to the force's parameters,
// written in the caller's atom order (perm: sorted -> caller):
//  gmol[20]  dE/dm of the molecular-frame record (internal order), the lab
//            potential vector pot taken back through the rotation of k_rotate
//            and the chirality flip (a transpose: lab = L m);
//  gpol[3]   dE/dalpha_k of the molecular-frame polarisabilities: at fixed
//            induced dipoles the energy holds 1/2 mu alpha_lab^-1 mu, whose
//            derivative is -1/2 eps E E^T (E the total field, alpha_lab E =
//            mu), projected on molecular axis k; plus the damping factor
//            (mean alpha)^(1/6) through the Thole terms (-hd / damp);
//  gthole    dE/dthole from ht.
// Sites without moments and polarisability (bare) get the charge component
// only unless all_sites: the other pair kernels leave their components out.
// hdef is reduced into *gdef: dE/d(default Thole width).
__global__ __launch_bounds__(128) void k_param_gradients(int n, int all_sites, const int* __restrict__ perm,
                                                          Box<double> box, const double4* __restrict__ pos,
                                                          const int4* __restrict__ axes, const double* __restrict__ mol,
                                                          const double* __restrict__ pol, const double* __restrict__ damp,
                                                          const int* __restrict__ chg, const acc_t* __restrict__ pot,
                                                          const double* __restrict__ Etot, const acc_t* __restrict__ hd,
                                                          const acc_t* __restrict__ hdef, const acc_t* __restrict__ ht,
                                                          double default_thole, double* gmol, double* gpol,
                                                          double* gthole, acc_t* gdef) {
    __shared__ double sh[128];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double hsum = 0.0;
    if (s < n) {
        const int a = perm[s];
        const int4 ax = axes[s];
        const bool bare = chg[s] != 0 && !all_sites;
        double v[kMP];
        for (int k = 0; k < kMP; k++) v[k] = (bare && k > 0) ? 0.0 : acc_get(pot[kMP * (size_t)s + k]);
        D3 e[3];
        frame_axes(s, pos, ax, box, e);
        double E[3][3];
        for (int k = 0; k < 3; k++) {
            E[k][0] = e[k].x;
            E[k][1] = e[k].y;
            E[k][2] = e[k].z;
        }
        // transpose of the rotation in k_rotate
        double g[kMP];
        g[0] = v[0];
        for (int k = 0; k < 3; k++) {
            double t = 0;
            for (int b = 0; b < 3; b++) t += E[k][b] * v[1 + b];
            g[1 + k] = t;
        }
        for (int c = 4; c < kMP; c++) g[c] = 0.0;
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) {
                double t = 0;
                for (int p = 0; p < 3; p++)
                    for (int q = p; q < 3; q++) t += E[k][p] * E[l][q] * v[4 + i2(p, q)];
                g[4 + i2(k, l)] += t;
            }
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
                for (int w = 0; w < 3; w++) {
                    double t = 0;
                    for (int p = 0; p < 3; p++)
                        for (int q = p; q < 3; q++)
                            for (int r = q; r < 3; r++) t += E[k][p] * E[l][q] * E[w][r] * v[10 + i3(p, q, r)];
                    g[10 + i3(k, l, w)] += t;
                }
        // chirality (ZThenX with a y atom): k_rotate flips these components
        if (ax.w == 0 && ax.z >= 0) {
            const D3 py = d3(pos[ax.z]);
            D3 ad = min_image(d3(pos[s]) - py, box), bd = min_image(d3(pos[ax.x]) - py, box),
               cd = min_image(d3(pos[ax.y]) - py, box);
            if (dot(cross(bd, cd), ad) < 0.0) {
                g[2] = -g[2];
                g[5] = -g[5];
                g[8] = -g[8];
            }
        }
        for (int k = 0; k < kMP; k++) gmol[kMP * (size_t)a + k] = g[k];
        // polarisabilities
        const double p0 = pol[3 * (size_t)s], p1 = pol[3 * (size_t)s + 1], p2 = pol[3 * (size_t)s + 2];
        const bool polar = p0 != 0.0 || p1 != 0.0 || p2 != 0.0;
        const D3 Et = mk(Etot[3 * (size_t)s], Etot[3 * (size_t)s + 1], Etot[3 * (size_t)s + 2]);
        const double pbar = (p0 + p1 + p2) / 3.0;
        const double dd = damp[s];
        // dE/ddamp = -hd / damp; ddamp/dalpha_k = damp / (18 mean alpha)
        const double ddamp = (dd > 0.0 && pbar > 0.0) ? -acc_get(hd[s]) / (18.0 * pbar) : 0.0;
        for (int k = 0; k < 3; k++) {
            const double ek = dot(e[k], Et);
            gpol[3 * (size_t)a + k] = (polar || all_sites) ? -0.5 * kElectric * ek * ek + ddamp : 0.0;
        }
        gthole[a] = acc_get(ht[s]);
        hsum = acc_get(hdef[s]);
    }
    sh[threadIdx.x] = hsum;
    __syncthreads();
    for (int off = 64; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0 && sh[0] != 0.0) acc_add(gdef, sh[0] / (2.0 * default_thole));
}

// ---------------------------------------------------------------------------
// torque -> forces on the frame atoms
// ---------------------------------------------------------------------------
// exact_threefold: ThreeFold frames by the exact chain rule of the frame
// (z = s/|s|, s = u + v + w the sum of the unit vectors to the z, x and y
// atoms; x = t/|t|, t = v - (v.z) z).  Moving an atom turns the frame by
// dtheta = z x dz + (y.dx) z, so with tau = -dE/dtheta the force on atom a
// is (dz/dr_a)^T (tau x z) + (tau.z) (dx/dr_a)^T y.  Without it, the
// reference's approximate projection (MPIDReferenceForce.cpp:2068-2097).
__global__ __launch_bounds__(128) void k_torque_to_force(int n, Box<double> box, const double4* __restrict__ pos,
                                                          const int4* __restrict__ axes,
                                                          const acc_t* __restrict__ torque, acc_t* force,
                                                          int exact_threefold) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int4 ax = axes[s];
    const int t = ax.w;
    if (t == 5 || ax.x < 0) return;
    const D3 tq = mk(acc_get(torque[3 * (size_t)s]), acc_get(torque[3 * (size_t)s + 1]),
                     acc_get(torque[3 * (size_t)s + 2]));
    const D3 pI = d3(pos[s]);
    const int iz = ax.x, ix = ax.y, iy = ax.z;
    if (t == 3 && exact_threefold && iy >= 0) {
        D3 u = min_image(d3(pos[iz]) - pI, box), v = min_image(d3(pos[ix]) - pI, box),
           w = min_image(d3(pos[iy]) - pI, box);
        const double nu = dnorm(u), nv = dnorm(v), nw = dnorm(w);
        D3 z = u + v + w;
        const double ns = dnorm(z);
        D3 x = v - z * dot(v, z);
        const double nt = dnorm(x);
        const D3 y = cross(z, x);
        const D3 c = y * (dot(tq, z) / nt);          // through x: on t, hence on v and on z
        const D3 bz = cross(tq, z) - c * dot(v, z);  // everything that moves with z
        const D3 gs = (bz - z * dot(z, bz)) * (1.0 / ns);
        const D3 gx = gs + c;
        const D3 FZ = (gs - u * dot(u, gs)) * (1.0 / nu);
        const D3 FX = (gx - v * dot(v, gx)) * (1.0 / nv);
        const D3 FY = (gs - w * dot(w, gs)) * (1.0 / nw);
        const D3 FI = (FZ + FX + FY) * -1.0;
        const int at[4] = {iz, ix, iy, s};
        const D3 fs[4] = {FZ, FX, FY, FI};
        for (int q = 0; q < 4; q++) {
            acc_add(&force[3 * (size_t)at[q] + 0], fs[q].x);
            acc_add(&force[3 * (size_t)at[q] + 1], fs[q].y);
            acc_add(&force[3 * (size_t)at[q] + 2], fs[q].z);
        }
        return;
    }
    D3 U = min_image(d3(pos[iz]) - pI, box);
    const double nU = dnorm(U);
    D3 V;
    if (t == 4) V = fabs(U.x) < 0.866 ? mk(1.0, 0.0, 0.0) : mk(0.0, 1.0, 0.0);
    else V = min_image(d3(pos[ix]) - pI, box);
    const double nV = dnorm(V);
    D3 W;
    if ((t == 2 || t == 3) && iy >= 0) W = min_image(d3(pos[iy]) - pI, box);
    else W = cross(U, V);
    const double nW = dnorm(W);
    D3 UV = cross(V, U), UW = cross(W, U), VW = cross(W, V);
    dnorm(UV);
    dnorm(UW);
    dnorm(VW);
    const double cUV = dot(U, V), sUV = sqrt(1.0 - cUV * cUV);
    const double cUW = dot(U, W), sUW = sqrt(1.0 - cUW * cUW);
    const double cVW = dot(V, W), sVW = sqrt(1.0 - cVW * cVW);
    const double dU = -dot(U, tq), dV = -dot(V, tq), dW = -dot(W, tq);
    D3 fZ = mk(0.0, 0.0, 0.0), fX = fZ, fY = fZ;
    if (t == 0 || t == 1) {
        double f1 = dV / (nU * sUV), f2 = dW / nU, f3 = -dU / (nV * sUV), f4 = 0.0;
        if (t == 1) {
            f2 *= 0.5;
            f4 = 0.5 * dW / nV;
        }
        fZ = UV * f1 + UW * f2;
        fX = UV * f3 + VW * f4;
    } else if (t == 2) {
        D3 Rv = V + W;
        D3 Sv = cross(U, Rv);
        dnorm(Rv);
        dnorm(Sv);
        D3 UR = cross(Rv, U), US = cross(Sv, U), VS = cross(Sv, V), WS = cross(Sv, W);
        dnorm(UR);
        dnorm(US);
        dnorm(VS);
        dnorm(WS);
        const double cUR = dot(U, Rv), sUR = sqrt(1.0 - cUR * cUR);
        const double cVS = dot(V, Sv), sVS = sqrt(1.0 - cVS * cVS);
        const double cWS = dot(W, Sv), sWS = sqrt(1.0 - cWS * cWS);
        D3 t1 = V - Sv * cVS, t2 = W - Sv * cWS;
        dnorm(t1);
        dnorm(t2);
        const double ut1c = dot(U, t1), ut1s = sqrt(1.0 - ut1c * ut1c);
        const double ut2c = dot(U, t2), ut2s = sqrt(1.0 - ut2c * ut2c);
        const double dR = -dot(Rv, tq), dS = -dot(Sv, tq);
        const double f1 = dR / (nU * sUR), f2 = dS / nU, f3 = dU / (nV * (ut1s + ut2s)), f4 = dU / (nW * (ut1s + ut2s));
        fZ = UR * f1 + US * f2;
        fX = (Sv * sVS - t1 * cVS) * f3;
        fY = (Sv * sWS - t2 * cWS) * f4;
    } else if (t == 3) {
        fZ = (UW * (dW / (nU * sUW)) + UV * (dV / (nU * sUV)) - UW * (dU / (nU * sUW)) - UV * (dU / (nU * sUV))) *
             (1.0 / 3.0);
        fX = (VW * (dW / (nV * sVW)) - UV * (dU / (nV * sUV)) - VW * (dV / (nV * sVW)) + UV * (dV / (nV * sUV))) *
             (1.0 / 3.0);
        fY = (UW * (-dU / (nW * sUW)) - VW * (dV / (nW * sVW)) + UW * (dW / (nW * sUW)) + VW * (dW / (nW * sVW))) *
             (1.0 / 3.0);
        if (iy < 0) fY = mk(0.0, 0.0, 0.0);
    } else if (t == 4) {
        fZ = UV * (dV / (nU * sUV)) + UW * (dW / nU);
    }
    const D3 fI = fZ + fX + fY;
    acc_add(&force[3 * (size_t)iz + 0], -fZ.x);
    acc_add(&force[3 * (size_t)iz + 1], -fZ.y);
    acc_add(&force[3 * (size_t)iz + 2], -fZ.z);
    if (t != 4) {
        acc_add(&force[3 * (size_t)ix + 0], -fX.x);
        acc_add(&force[3 * (size_t)ix + 1], -fX.y);
        acc_add(&force[3 * (size_t)ix + 2], -fX.z);
    }
    if ((t == 2 || t == 3) && iy >= 0) {
        acc_add(&force[3 * (size_t)iy + 0], -fY.x);
        acc_add(&force[3 * (size_t)iy + 1], -fY.y);
        acc_add(&force[3 * (size_t)iy + 2], -fY.z);
    }
    acc_add(&force[3 * (size_t)s + 0], fI.x);
    acc_add(&force[3 * (size_t)s + 1], fI.y);
    acc_add(&force[3 * (size_t)s + 2], fI.z);
}

__global__ void k_output_forces(int n, const int* __restrict__ perm, const acc_t* __restrict__ force, double* out,
                                int accumulate) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int a = perm[s];
    for (int k = 0; k < 3; k++) {
        const double v = acc_get(force[3 * (size_t)s + k]);
        if (accumulate) out[3 * (size_t)a + k] += v;
        else out[3 * (size_t)a + k] = v;
    }
}

__global__ void k_unpermute3(int n, const int* __restrict__ perm, const double* __restrict__ in, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int a = perm[s];
    for (int k = 0; k < 3; k++) out[3 * (size_t)a + k] = in[3 * (size_t)s + k];
}

template <class R>
__global__ void k_field_sum(int n, const acc_t* __restrict__ acc, const R* __restrict__ erec, const R* __restrict__ mu,
                            double selfc, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    for (int k = 0; k < 3; k++) {
        const size_t id = 3 * (size_t)s + k;
        out[id] = acc_get(acc[id]) + (double)erec[id] + selfc * (double)mu[id];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t sort_temp_bytes(int n) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr, (int*)nullptr,
                              (size_t)n, 0, 30);
    return bytes;
}
int launch_sort_pairs(hipStream_t s, void* temp, size_t bytes, unsigned* kin, unsigned* kout, int* vin, int* vout,
                      int n) {
    return (int)rocprim::radix_sort_pairs(temp, bytes, kin, kout, vin, vout, (size_t)n, 0, 30, s);
}
template <class R>
void launch_field_sum(hipStream_t s, int n, const acc_t* acc, const R* erec, const R* mu, double selfc, double* out) {
    hipLaunchKernelGGL(k_field_sum<R>, dim3((n + 255) / 256), dim3(256), 0, s, n, acc, erec, mu, selfc, out);
}

static inline dim3 blocks_for(size_t n, int t) { return dim3((unsigned)((n + t - 1) / t)); }

void launch_morton_keys(hipStream_t s, int n, const double* pos, Box<double> box, unsigned* keys, int* vals) {
    hipLaunchKernelGGL(k_morton_keys, blocks_for(n, 256), dim3(256), 0, s, n, pos, box, keys, vals);
}
void launch_invert_perm(hipStream_t s, int n, const int* perm, int* inv) {
    hipLaunchKernelGGL(k_invert_perm, blocks_for(n, 256), dim3(256), 0, s, n, perm, inv);
}
void launch_gather_sorted(hipStream_t s, int n, const int* perm, const int* inv, const double* pos, Box<double> box,
                          const double* mol, const int4* axes, const double* thole, const double* pol,
                          const int* partners, const int* npart, int pstride, double4* spos, double* smol,
                          int4* saxes, double* sthole, double* spol, int* spartners, int* snpart) {
    hipLaunchKernelGGL(k_gather_sorted, blocks_for(n, 128), dim3(128), 0, s, n, perm, inv, pos, box, mol, axes, thole,
                       pol, partners, npart, pstride, spos, smol, saxes, sthole, spol, spartners, snpart);
}
void launch_rotate(hipStream_t s, int n, Box<double> box, const double4* pos, const int4* axes, const double* mol,
                   const double* pol, double* lab, double* labpol, double* sqrtpol, double* damp, int* aniso,
                   int* chg) {
    hipLaunchKernelGGL(k_rotate, blocks_for(n, 128), dim3(128), 0, s, n, box, pos, axes, mol, pol, lab, labpol,
                       sqrtpol, damp, aniso, chg);
}
template <class R>
void launch_convert(hipStream_t s, int n, const double4* pos, const double* lab, const double* thole,
                    const double* damp, R4<R>* rpos, R* rlab, R2<R>* rtd) {
    hipLaunchKernelGGL(k_convert<R>, blocks_for(n, 256), dim3(256), 0, s, n, pos, lab, thole, damp, rpos, rlab, rtd);
}
template <class R>
void launch_recip_potential(hipStream_t s, int n, int pme, const double* phi, const R* phidp, const double* lab,
                            const double* mu, FracT<double> A, double alpha, acc_t* pot) {
    hipLaunchKernelGGL(k_recip_potential<R>, blocks_for(n, 128), dim3(128), 0, s, n, pme, phi, phidp, lab, mu, A,
                       alpha, pot);
}
template <class R>
void launch_total_field(hipStream_t s, int n, int mutual, const double* E0, const acc_t* acc, const R* phidp,
                        FracT<double> A, const double* mu, double selfc, double* out) {
    hipLaunchKernelGGL(k_total_field<R>, blocks_for(n, 256), dim3(256), 0, s, n, mutual, E0, acc, phidp, A, mu, selfc,
                       out);
}
void launch_param_gradients(hipStream_t s, int n, int all_sites, const int* perm, Box<double> box, const double4* pos,
                            const int4* axes, const double* mol, const double* pol, const double* damp,
                            const int* chg, const acc_t* pot, const double* Etot, const acc_t* hd,
                            const acc_t* hdef, const acc_t* ht, double default_thole, double* gmol, double* gpol,
                            double* gthole, acc_t* gdef) {
    hipLaunchKernelGGL(k_param_gradients, blocks_for(n, 128), dim3(128), 0, s, n, all_sites, perm, box, pos, axes, mol,
                       pol, damp, chg, pot, Etot, hd, hdef, ht, default_thole, gmol, gpol, gthole, gdef);
}
void launch_torque_to_force(hipStream_t s, int n, Box<double> box, const double4* pos, const int4* axes,
                            const acc_t* torque, acc_t* force, int exact_threefold) {
    hipLaunchKernelGGL(k_torque_to_force, blocks_for(n, 128), dim3(128), 0, s, n, box, pos, axes, torque, force,
                       exact_threefold);
}
void launch_output_forces(hipStream_t s, int n, const int* perm, const acc_t* force, double* out, int accumulate) {
    hipLaunchKernelGGL(k_output_forces, blocks_for(n, 256), dim3(256), 0, s, n, perm, force, out, accumulate);
}
void launch_unpermute3(hipStream_t s, int n, const int* perm, const double* in, double* out) {
    hipLaunchKernelGGL(k_unpermute3, blocks_for(n, 256), dim3(256), 0, s, n, perm, in, out);
}
void launch_zero_acc(hipStream_t s, acc_t* p, size_t count, const double* skip) {
    int b = (int)std::min<size_t>((count + 255) / 256, 4096);
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_zero_acc, dim3(b), dim3(256), 0, s, p, count, skip);
}
template <class R> void launch_fill(hipStream_t s, R* p, size_t count, R v) {
    int b = (int)std::min<size_t>((count + 255) / 256, 8192);
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_fill<R>, dim3(b), dim3(256), 0, s, p, count, v);
}
template <class R>
void launch_fixed_field_finish(hipStream_t s, int n, const R* phi, const double* lab, const acc_t* acc,
                               FracT<double> A, double selfc, double* E0) {
    hipLaunchKernelGGL(k_fixed_field_finish<R>, blocks_for(n, 256), dim3(256), 0, s, n, phi, lab, acc, A, selfc, E0);
}
void launch_solver_init(hipStream_t s, int n, const double* S, const double* E0, double* b, double* y) {
    hipLaunchKernelGGL(k_solver_init, blocks_for(n, 256), dim3(256), 0, s, n, S, E0, b, y);
}
template <class R> void launch_apply_S(hipStream_t s, int n, const double* S, const double* y, R* mu, double* mud) {
    hipLaunchKernelGGL(k_apply_S<R>, blocks_for(n, 256), dim3(256), 0, s, n, S, y, mu, mud);
}
template <class R>
void launch_matvec_finish(hipStream_t s, int n, const double* S, const acc_t* acc, const R* erecip, const R* mu,
                          double selfc, const double* p, double* Ap) {
    hipLaunchKernelGGL(k_matvec_finish<R>, blocks_for(n, 256), dim3(256), 0, s, n, S, acc, erecip, mu, selfc, p, Ap);
}
void launch_dot(hipStream_t s, int m, const double* a, const double* b, double* partial, double* result) {
    hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, s, m, a, b, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedThreads), 0, s, partial, 1, result, nullptr);
}
void launch_residual_norm(hipStream_t s, int n, const double* S, const double* r, double* partial, double* result) {
    hipLaunchKernelGGL(k_residual_norm, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, S, r, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedThreads), 0, s, partial, 2, result, nullptr);
}
// One CG iteration after the operator's field is complete (scal slots:
// [0] rr, [1] pAp, [2] rr_new, [3] ss, [7] beta): Ap and pAp, x += a p,
// r -= a Ap with their norms, the iteration's check, then the next direction
// p and its mu = S p.  Deterministic reductions: per-block partials, then
// one block in block order.  Every kernel returns at once when the solve
// has stopped before the iteration.
template <class R>
void launch_cg_iteration(hipStream_t s, int n, const double* S, const acc_t* acc, const R* erec, R* mu, double selfc,
                         double* p, double* Ap, double* x, double* r, double* partial, double* scal, double* hcheck,
                         double stop_eps) {
    hipLaunchKernelGGL(k_matvec_finish_dot<R>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, S, acc, erec, mu, selfc,
                       p, Ap, partial, scal);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedThreads), 0, s, partial, 1, scal + 1, scal + kCgStop);
    hipLaunchKernelGGL(k_cg_step, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, S, scal, p, Ap, x, r, partial);
    hipLaunchKernelGGL(k_cg_reduce_beta, dim3(1), dim3(kRedThreads), 0, s, partial, scal, hcheck, n, stop_eps);
    hipLaunchKernelGGL(k_cg_direction_apply<R>, blocks_for(n, 256), dim3(256), 0, s, n, S, scal, r, p, mu);
}
template <class R>
void launch_extrap_order(hipStream_t s, int n, const acc_t* field, const acc_t* gacc, const R* phidp,
                         FracT<double> A, double selfc, const double* labpol, const double* mu, double* E, double* G,
                         double* mu_next, R* mu_next_r) {
    hipLaunchKernelGGL(k_extrap_order<R>, blocks_for(n, 256), dim3(256), 0, s, n, field, gacc, phidp, A, selfc,
                       labpol, mu, E, G, mu_next, mu_next_r);
}
template <class R>
void launch_extrap_combine(hipStream_t s, int n, const ExtrapCoefficients& cp, const double* mus, double* mud,
                           R* mu_r) {
    hipLaunchKernelGGL(k_extrap_combine<R>, blocks_for(3 * n, 256), dim3(256), 0, s, n, cp, mus, mud, mu_r);
}
void launch_extrap_response(hipStream_t s, int n, const ExtrapCoefficients& cp, const double* mus, const double* Es,
                            const double* Gs, const int* aniso, acc_t* force, acc_t* torque) {
    hipLaunchKernelGGL(k_extrap_response, blocks_for(n, 256), dim3(256), 0, s, n, cp, mus, Es, Gs, aniso, force,
                       torque);
}
void launch_potential_points(hipStream_t s, int n, int npts, const double4* pos, const double* lab, const double* mu,
                             Box<double> box, const double* points, double* out) {
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_potential_points, dim3((npts + kPotThreads - 1) / kPotThreads), dim3(kPotThreads), 0, s, n,
                       npts, pos, lab, mu, box, points, out);
}

// Staging memory of the allocation-free potential getter: the points
// (3 doubles each) at [0, 3 kNoAllocMaxPoints), their potentials after them.
// Every GPU the library runs on holds its own copy until the process ends.
// Host code never passes the name itself to a kernel (on the host it is a
// stand-in, not a device address): it uses the address that
// noalloc_scratch_address() looked up for the current device.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

hipError_t noalloc_scratch_address(double** out) {
    static std::mutex lock;
    static std::map<int, double*> addr;  // device -> address of that device's copy
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(lock);
    auto it = addr.find(dev);
    if (it != addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    double* p = nullptr;
    e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = p;
    *out = p;
    return hipSuccess;
}
void launch_multipole_moments(hipStream_t s, int n, const int* perm, const double* pos, const double* masses,
                              const double* lab, const double* mu, double* out) {
    hipLaunchKernelGGL(k_multipole_moments, dim3(1), dim3(kMomThreads), 0, s, n, perm, pos, masses, lab, mu, out);
}
void launch_cg_begin(hipStream_t s, int n, double* scal, double* hcheck, double stop_eps) {
    hipLaunchKernelGGL(k_cg_begin, dim3(1), dim3(64), 0, s, scal, hcheck, n, stop_eps);
}
void launch_axpy(hipStream_t s, int m, double a, const double* x, double* y) {
    int b = std::min((m + 255) / 256, 4096);
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_axpy, dim3(b), dim3(256), 0, s, m, a, x, y);
}
void launch_sub(hipStream_t s, int m, const double* a, const double* b_, double* out) {
    int b = std::min((m + 255) / 256, 4096);
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_sub, dim3(b), dim3(256), 0, s, m, a, b_, out);
}
template <class R>
void launch_recip_atoms(hipStream_t s, int n, int mutual, const double* phi, const R* phidp, const double* lab,
                        const double* mu, const int* aniso, FracT<double> A, double alpha, acc_t* force,
                        acc_t* torque, acc_t* energy) {
    hipLaunchKernelGGL(k_recip_atoms<R>, blocks_for(n, 128), dim3(128), 0, s, n, mutual, phi, phidp, lab, mu, aniso,
                       A, alpha, force, torque, energy);
}

#define MPID_INSTANTIATE(R)                                                                                         \
    template void launch_convert<R>(hipStream_t, int, const double4*, const double*, const double*, const double*,  \
                                    R4<R>*, R*, R2<R>*);                                                            \
    template void launch_fill<R>(hipStream_t, R*, size_t, R);                                                        \
    template void launch_fixed_field_finish<R>(hipStream_t, int, const R*, const double*, const acc_t*, FracT<double>, \
                                               double, double*);                                                    \
    template void launch_apply_S<R>(hipStream_t, int, const double*, const double*, R*, double*);                   \
    template void launch_matvec_finish<R>(hipStream_t, int, const double*, const acc_t*, const R*, const R*, double, \
                                          const double*, double*);                                                  \
    template void launch_extrap_order<R>(hipStream_t, int, const acc_t*, const acc_t*, const R*, FracT<double>, double,   \
                                         const double*, const double*, double*, double*, double*, R*);              \
    template void launch_extrap_combine<R>(hipStream_t, int, const ExtrapCoefficients&, const double*, double*, R*); \
    template void launch_cg_iteration<R>(hipStream_t, int, const double*, const acc_t*, const R*, R*, double,       \
                                         double*, double*, double*, double*, double*, double*, double*, double);    \
    template void launch_recip_atoms<R>(hipStream_t, int, int, const double*, const R*, const double*, const double*,\
                                        const int*, FracT<double>, double, acc_t*, acc_t*, acc_t*);                      \
    template void launch_field_sum<R>(hipStream_t, int, const acc_t*, const R*, const R*, double, double*);                    \
    template void launch_recip_potential<R>(hipStream_t, int, int, const double*, const R*, const double*,            \
                                            const double*, FracT<double>, double, acc_t*);                                 \
    template void launch_total_field<R>(hipStream_t, int, int, const double*, const acc_t*, const R*, FracT<double>,    \
                                        const double*, double, double*);
MPID_INSTANTIATE(float)
MPID_INSTANTIATE(double)
#undef MPID_INSTANTIATE

}  // namespace mpid
