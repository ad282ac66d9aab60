// Real-space electrostatics (energy, forces, torques) of the MPIDForce PME
// path on gfx950, and the covalent-pair kernels.
//
// Atoms are sorted along a Morton curve and cut into blocks of 64 (one
// wavefront).  A tile is a pair of blocks whose bounding boxes come within the
// cutoff.  One wavefront owns one tile: lane t holds atom i = I*64+t in
// registers, the 64 atoms of block J sit in LDS, and in step s lane t meets
// j = J*64 + (t+s)%64.  In every step the 64 lanes touch 64 distinct j, so the
// j-side accumulators live in LDS without atomics; both sides are flushed to
// the fixed-point global accumulators once per tile.
//
// The pair formulas are in pair_math.h (MPIDReferenceForce.cpp:4335-4920).
#include "pair_math.h"

namespace mpid {

// ---------------------------------------------------------------------------
// blocks and tiles
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(64) void k_block_bounds(int n, const R4<R>* __restrict__ pos, float4* bmin,
                                                      float4* bmax) {
    int b = blockIdx.x, lane = threadIdx.x;
    int i = b * kTile + lane;
    float lo[3], hi[3];
    if (i < n) {
        R4<R> p = pos[i];
        lo[0] = hi[0] = (float)p.x;
        lo[1] = hi[1] = (float)p.y;
        lo[2] = hi[2] = (float)p.z;
    } else {
        R4<R> p = pos[b * kTile];
        lo[0] = hi[0] = (float)p.x;
        lo[1] = hi[1] = (float)p.y;
        lo[2] = hi[2] = (float)p.z;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
        }
    }
    if (lane == 0) {
        bmin[b] = make_float4(lo[0], lo[1], lo[2], 0.f);
        bmax[b] = make_float4(hi[0], hi[1], hi[2], 0.f);
    }
}

__global__ void k_build_tiles(int nb, const float4* __restrict__ bmin, const float4* __restrict__ bmax,
                              Box<float> box, float cutoff, int2* tiles, int* count, int capacity) {
    int I = blockIdx.y;
    int J = blockIdx.x * blockDim.x + threadIdx.x;
    if (J >= nb || J < I) return;
    float4 a0 = bmin[I], a1 = bmax[I], b0 = bmin[J], b1 = bmax[J];
    float ca[3] = {0.5f * (a0.x + a1.x), 0.5f * (a0.y + a1.y), 0.5f * (a0.z + a1.z)};
    float ha[3] = {0.5f * (a1.x - a0.x), 0.5f * (a1.y - a0.y), 0.5f * (a1.z - a0.z)};
    float cb[3] = {0.5f * (b0.x + b1.x), 0.5f * (b0.y + b1.y), 0.5f * (b0.z + b1.z)};
    float hb[3] = {0.5f * (b1.x - b0.x), 0.5f * (b1.y - b0.y), 0.5f * (b1.z - b0.z)};
    float d2 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float d = cb[k] - ca[k];
        d -= box.L[k] * floorf(d * box.inv[k] + 0.5f);
        float g = fmaxf(0.f, fabsf(d) - ha[k] - hb[k]);
        d2 += g * g;
    }
    // a little slack absorbs float rounding of the bounds
    float rc = cutoff * 1.0001f + 1e-5f;
    if (d2 <= rc * rc) {
        int slot = atomicAdd(count, 1);
        if (slot < capacity) tiles[slot] = make_int2(I, J);
    }
}

template <class R>
__global__ __launch_bounds__(64) void k_electrostatics(const int2* __restrict__ tiles, PairAtoms<R> A,
                                                        const R* __restrict__ mu, PairParams<R> P,
                                                        acc_t* __restrict__ force, acc_t* __restrict__ torque,
                                                        acc_t* __restrict__ energy) {
    __shared__ R4<R> sp[kTile];
    __shared__ R sm[kTile][kMP];
    __shared__ R smu[kTile][3];
    __shared__ R2<R> std_[kTile];
    __shared__ int san[kTile];
    __shared__ R sacc[kTile][6];
    const int lane = threadIdx.x;
    const int2 t = tiles[blockIdx.x];
    const int I = t.x, J = t.y;
    const int n = P.n;
    {
        int j = J * kTile + lane;
        int jc = j < n ? j : n - 1;
        sp[lane] = A.pos[jc];
#pragma unroll
        for (int k = 0; k < kMP; k++) sm[lane][k] = A.lab[(size_t)jc * kMP + k];
        smu[lane][0] = mu[3 * (size_t)jc + 0];
        smu[lane][1] = mu[3 * (size_t)jc + 1];
        smu[lane][2] = mu[3 * (size_t)jc + 2];
        std_[lane] = A.td[jc];
        san[lane] = A.aniso[jc];
#pragma unroll
        for (int k = 0; k < 6; k++) sacc[lane][k] = R(0);
    }
    const int i = I * kTile + lane;
    const bool iv = i < n;
    const int ic = iv ? i : n - 1;
    const R4<R> pi = A.pos[ic];
    R mi[kMP];
#pragma unroll
    for (int k = 0; k < kMP; k++) mi[k] = A.lab[(size_t)ic * kMP + k];
    const V3<R> ui = mk(mu[3 * (size_t)ic], mu[3 * (size_t)ic + 1], mu[3 * (size_t)ic + 2]);
    const R2<R> tdi = A.td[ic];
    const bool ani = A.aniso[ic] != 0;
    const unsigned long long pm = partner_mask(A.partners + (size_t)ic * kMaxPartners, iv ? A.npart[ic] : 0, J);
    R fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
    // energies summed in FP64: the real-space total is a large number that the
    // reciprocal and self terms cancel almost entirely
    double en = 0.0;
    __syncthreads();
    for (int s = 0; s < kTile; s++) {
        const int jj = (lane + s) & (kTile - 1);
        const int j = J * kTile + jj;
        if (!iv || j >= n || (I == J && jj <= lane) || ((pm >> jj) & 1ull)) continue;
        const R4<R> pj = sp[jj];
        V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        const R r2 = dot(dr, dr);
        if (r2 > P.rc2) continue;
        ElecOut<R> o = elec_pair<R>(dr, r2, mi, sm[jj], ui, mk(smu[jj][0], smu[jj][1], smu[jj][2]), tdi, std_[jj],
                                    ani, san[jj] != 0, R(1), R(1), P);
        en += (double)o.energy;
        fx -= o.f.x;
        fy -= o.f.y;
        fz -= o.f.z;
        tx += o.ti.x;
        ty += o.ti.y;
        tz += o.ti.z;
        sacc[jj][0] += o.f.x;
        sacc[jj][1] += o.f.y;
        sacc[jj][2] += o.f.z;
        sacc[jj][3] += o.tj.x;
        sacc[jj][4] += o.tj.y;
        sacc[jj][5] += o.tj.z;
    }
    __syncthreads();
    if (iv) {
        acc_add(&force[3 * (size_t)i + 0], fx);
        acc_add(&force[3 * (size_t)i + 1], fy);
        acc_add(&force[3 * (size_t)i + 2], fz);
        acc_add(&torque[3 * (size_t)i + 0], tx);
        acc_add(&torque[3 * (size_t)i + 1], ty);
        acc_add(&torque[3 * (size_t)i + 2], tz);
    }
    const int j = J * kTile + lane;
    if (j < n) {
        acc_add(&force[3 * (size_t)j + 0], sacc[lane][0]);
        acc_add(&force[3 * (size_t)j + 1], sacc[lane][1]);
        acc_add(&force[3 * (size_t)j + 2], sacc[lane][2]);
        acc_add(&torque[3 * (size_t)j + 0], sacc[lane][3]);
        acc_add(&torque[3 * (size_t)j + 1], sacc[lane][4]);
        acc_add(&torque[3 * (size_t)j + 2], sacc[lane][5]);
    }
    // wavefront sum of the pair energies, one fixed-point add per tile
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) en += __shfl_xor(en, off);
    if (lane == 0) acc_add(energy, en);
}

// ---------------------------------------------------------------------------
// Covalent pairs.  The pairs on the scaled covalent lists (1-2, 1-3 and 1-4
// partners) are few, O(N), and carry the largest terms of the real-space sum:
// at 0.1 nm the screened and unscreened Coulomb parts cancel to a few percent,
// and the same geometry repeats in every molecule, so single-precision
// rounding would add up coherently.  The tile kernels skip them; these
// kernels evaluate each once (thread = atom i, partners j > i), in the
// precision they are instantiated with (FP64 for the permanent field and the
// final electrostatics in both precision modes).
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void k_cov_fixed_field(PairAtoms<R> A, PairParams<R> P, acc_t* __restrict__ field) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int np = A.npart[i];
    if (np == 0) return;
    const R4<R> pi = A.pos[i];
    const R* mi = A.lab + (size_t)i * kMP;
    const R2<R> tdi = A.td[i];
    R ex = 0, ey = 0, ez = 0;
    for (int k = 0; k < np; k++) {
        const int e = A.partners[(size_t)i * kMaxPartners + k];
        const int j = e >> 2;
        if (j <= i) continue;
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        const R r2 = dot(dr, dr);
        if (r2 > P.rc2) continue;
        V3<R> fi, fj;
        fixed_field_pair<R>(dr, r2, mi, A.lab + (size_t)j * kMP, tdi, A.td[j], partner_scale<R>(e, P.scale14), P, fi,
                            fj);
        ex += fi.x;
        ey += fi.y;
        ez += fi.z;
        acc_add(&field[3 * (size_t)j + 0], fj.x);
        acc_add(&field[3 * (size_t)j + 1], fj.y);
        acc_add(&field[3 * (size_t)j + 2], fj.z);
    }
    acc_add(&field[3 * (size_t)i + 0], ex);
    acc_add(&field[3 * (size_t)i + 1], ey);
    acc_add(&field[3 * (size_t)i + 2], ez);
}

template <class R>
__global__ __launch_bounds__(256) void k_cov_induced_field(PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                                                           acc_t* __restrict__ field) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int np = A.npart[i];
    if (np == 0) return;
    const R4<R> pi = A.pos[i];
    const V3<R> ui = mk(mu[3 * (size_t)i], mu[3 * (size_t)i + 1], mu[3 * (size_t)i + 2]);
    const R2<R> tdi = A.td[i];
    R ex = 0, ey = 0, ez = 0;
    for (int k = 0; k < np; k++) {
        const int e = A.partners[(size_t)i * kMaxPartners + k];
        const int j = e >> 2;
        if (j <= i) continue;
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        const R r2 = dot(dr, dr);
        if (r2 > P.rc2) continue;
        V3<R> fi, fj;
        induced_field_pair<R>(dr, r2, ui, mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]), tdi,
                              A.td[j], partner_scale<R>(e, P.scale14), P, fi, fj);
        ex += fi.x;
        ey += fi.y;
        ez += fi.z;
        acc_add(&field[3 * (size_t)j + 0], fj.x);
        acc_add(&field[3 * (size_t)j + 1], fj.y);
        acc_add(&field[3 * (size_t)j + 2], fj.z);
    }
    acc_add(&field[3 * (size_t)i + 0], ex);
    acc_add(&field[3 * (size_t)i + 1], ey);
    acc_add(&field[3 * (size_t)i + 2], ez);
}

template <class R>
__global__ __launch_bounds__(128) void k_cov_electrostatics(PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                                                            acc_t* __restrict__ force, acc_t* __restrict__ torque,
                                                            acc_t* __restrict__ energy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double en = 0.0;
    if (i < P.n && A.npart[i] > 0) {
        const int np = A.npart[i];
        const R4<R> pi = A.pos[i];
        const R* mi = A.lab + (size_t)i * kMP;
        const V3<R> ui = mk(mu[3 * (size_t)i], mu[3 * (size_t)i + 1], mu[3 * (size_t)i + 2]);
        const R2<R> tdi = A.td[i];
        const bool ani = A.aniso[i] != 0;
        R fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
        bool any = false;
        for (int k = 0; k < np; k++) {
            const int e = A.partners[(size_t)i * kMaxPartners + k];
            const int j = e >> 2;
            if (j <= i) continue;
            const R4<R> pj = A.pos[j];
            const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
            const R r2 = dot(dr, dr);
            if (r2 > P.rc2) continue;
            const R sc = partner_scale<R>(e, P.scale14);
            const ElecOut<R> o =
                elec_pair<R>(dr, r2, mi, A.lab + (size_t)j * kMP, ui,
                             mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]), tdi, A.td[j], ani,
                             A.aniso[j] != 0, sc, sc, P);
            any = true;
            en += (double)o.energy;
            fx -= o.f.x;
            fy -= o.f.y;
            fz -= o.f.z;
            tx += o.ti.x;
            ty += o.ti.y;
            tz += o.ti.z;
            acc_add(&force[3 * (size_t)j + 0], o.f.x);
            acc_add(&force[3 * (size_t)j + 1], o.f.y);
            acc_add(&force[3 * (size_t)j + 2], o.f.z);
            acc_add(&torque[3 * (size_t)j + 0], o.tj.x);
            acc_add(&torque[3 * (size_t)j + 1], o.tj.y);
            acc_add(&torque[3 * (size_t)j + 2], o.tj.z);
        }
        if (any) {
            acc_add(&force[3 * (size_t)i + 0], fx);
            acc_add(&force[3 * (size_t)i + 1], fy);
            acc_add(&force[3 * (size_t)i + 2], fz);
            acc_add(&torque[3 * (size_t)i + 0], tx);
            acc_add(&torque[3 * (size_t)i + 1], ty);
            acc_add(&torque[3 * (size_t)i + 2], tz);
        }
    }
    // every lane reaches the reduction
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) en += __shfl_xor(en, off);
    if ((threadIdx.x & 63) == 0 && en != 0.0) acc_add(energy, en);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class R>
void launch_block_bounds(hipStream_t s, int n, const R4<R>* pos, float4* bmin, float4* bmax) {
    int nb = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_block_bounds<R>, dim3(nb), dim3(kTile), 0, s, n, pos, bmin, bmax);
}

void launch_build_tiles(hipStream_t s, int nb, const float4* bmin, const float4* bmax, Box<float> box, float cutoff,
                        int2* tiles, int* count, int capacity) {
    dim3 grid((nb + 63) / 64, nb);
    hipLaunchKernelGGL(k_build_tiles, grid, dim3(64), 0, s, nb, bmin, bmax, box, cutoff, tiles, count, capacity);
}

template <class R>
void launch_electrostatics_pairs(hipStream_t s, int ntiles, const int2* tiles, PairAtoms<R> a, const R* mu,
                                 PairParams<R> p, acc_t* force, acc_t* torque, acc_t* energy) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_electrostatics<R>, dim3(ntiles), dim3(kTile), 0, s, tiles, a, mu, p, force, torque, energy);
}

template <class R> void launch_cov_fixed_field(hipStream_t s, int n, PairAtoms<R> a, PairParams<R> p, acc_t* field) {
    hipLaunchKernelGGL(k_cov_fixed_field<R>, dim3((n + 255) / 256), dim3(256), 0, s, a, p, field);
}
template <class R>
void launch_cov_induced_field(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, acc_t* field) {
    hipLaunchKernelGGL(k_cov_induced_field<R>, dim3((n + 255) / 256), dim3(256), 0, s, a, mu, p, field);
}
template <class R>
void launch_cov_electrostatics(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, acc_t* force,
                               acc_t* torque, acc_t* energy) {
    hipLaunchKernelGGL(k_cov_electrostatics<R>, dim3((n + 127) / 128), dim3(128), 0, s, a, mu, p, force, torque,
                       energy);
}

#define MPID_INSTANTIATE(R)                                                                                        \
    template void launch_block_bounds<R>(hipStream_t, int, const R4<R>*, float4*, float4*);                         \
    template void launch_electrostatics_pairs<R>(hipStream_t, int, const int2*, PairAtoms<R>, const R*,            \
                                                 PairParams<R>, acc_t*, acc_t*, acc_t*);             \
    template void launch_cov_fixed_field<R>(hipStream_t, int, PairAtoms<R>, PairParams<R>, acc_t*);                \
    template void launch_cov_induced_field<R>(hipStream_t, int, PairAtoms<R>, const R*, PairParams<R>, acc_t*);    \
    template void launch_cov_electrostatics<R>(hipStream_t, int, PairAtoms<R>, const R*, PairParams<R>, acc_t*,    \
                                               acc_t*, acc_t*);
MPID_INSTANTIATE(float)
MPID_INSTANTIATE(double)
#undef MPID_INSTANTIATE

}  // namespace mpid
