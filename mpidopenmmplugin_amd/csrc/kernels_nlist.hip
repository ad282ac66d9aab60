// Neighbour lists and the list-driven field kernels of the MPIDForce PME path.
//
// The solver applies the real-space dipole field operator once per iteration,
// so the pairs within the cutoff are found once per evaluation and kept as a
// full (both directions) per-atom list.  Each atom then sums only its own
// field: no atomics per pair, no second pass over pair slots that fail the
// cutoff, and the same list serves the permanent-field pass.
//
// Construction, all in the Morton-sorted order of the atoms:
//   1. block rows: for every 64-atom block I, the blocks J whose bounding
//      boxes come within the cutoff (both J < I and J >= I), in ascending J;
//   2. atom lists: lane i of block I tests the atoms of each J in its row and
//      appends those within the cutoff, skipping itself and its scaled
//      covalent partners (those pairs go to the covalent-pair kernels).
// Both steps are deterministic, so the summation order of every field is.
// A second list keeps only the polarisable neighbours: an atom without
// polarisability carries no induced dipole (mu = S y with S = 0) and its
// row of the solver's operator is multiplied by S = 0, so the solver's field
// passes run over the polarisable atoms and their polarisable neighbours only.
//
// List layout: entry k of atom i sits at ((k / S) * n + i) * S + k % S with
// S = kNlSplit, so the S threads that share an atom read S adjacent words and
// a wavefront reads 64 consecutive words per step.
#include "pair_math.h"

namespace mpid {

__device__ __forceinline__ size_t nl_index(int i, int k, int n) {
    return ((size_t)(k / kNlSplit) * n + i) * kNlSplit + (k % kNlSplit);
}

// Bounding boxes of blocks I and J within the cutoff (minimum image between
// the box centres, with a little slack for the float rounding of the bounds).
__device__ __forceinline__ bool blocks_within(float4 a0, float4 a1, float4 b0, float4 b1, const Box<float>& box,
                                              float cutoff) {
    const float ca[3] = {0.5f * (a0.x + a1.x), 0.5f * (a0.y + a1.y), 0.5f * (a0.z + a1.z)};
    const float ha[3] = {0.5f * (a1.x - a0.x), 0.5f * (a1.y - a0.y), 0.5f * (a1.z - a0.z)};
    const float cb[3] = {0.5f * (b0.x + b1.x), 0.5f * (b0.y + b1.y), 0.5f * (b0.z + b1.z)};
    const float hb[3] = {0.5f * (b1.x - b0.x), 0.5f * (b1.y - b0.y), 0.5f * (b1.z - b0.z)};
    float d2 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float d = cb[k] - ca[k];
        d -= box.L[k] * floorf(d * box.inv[k] + 0.5f);
        const float g = fmaxf(0.f, fabsf(d) - ha[k] - hb[k]);
        d2 += g * g;
    }
    const float rc = cutoff * 1.0001f + 1e-5f;
    return d2 <= rc * rc;
}

// One wavefront per block I.  Count mode (rows == nullptr) writes counts[I];
// fill mode writes the row at rows[start[I] ...] in ascending J, and I into
// owner[] at the same positions.
__global__ __launch_bounds__(64) void k_block_rows(int nb, const float4* __restrict__ bmin,
                                                   const float4* __restrict__ bmax, Box<float> box, float cutoff,
                                                   const int* __restrict__ start, int* __restrict__ counts,
                                                   int* __restrict__ rows, int* __restrict__ owner) {
    const int I = blockIdx.x, lane = threadIdx.x;
    const float4 a0 = bmin[I], a1 = bmax[I];
    const int base = rows ? start[I] : 0;
    int c = 0;
    for (int J0 = 0; J0 < nb; J0 += kTile) {
        const int J = J0 + lane;
        const bool hit = J < nb && blocks_within(a0, a1, bmin[J], bmax[J], box, cutoff);
        const unsigned long long m = __ballot(hit);
        if (hit && rows) {
            const int at = base + c + __popcll(m & ((1ull << lane) - 1ull));
            rows[at] = J;
            owner[at] = I;
        }
        c += __popcll(m);
    }
    if (lane == 0 && !rows) counts[I] = c;
}

// start[k] = counts[0] + ... + counts[k-1] for k = 0..m (one workgroup).
__global__ __launch_bounds__(1024) void k_exclusive_scan(int m, const int* __restrict__ counts, int* __restrict__ start) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (m + 1023) / 1024;
    const int b = t * per, e = min(b + per, m);
    int s = 0;
    for (int k = b; k < e; k++) s += counts[k];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t ? part[t - 1] : 0;
    for (int k = b; k < e; k++) {
        start[k] = run;
        run += counts[k];
    }
    if (t == 1023) start[m] = part[1023];
}

// Lane i of block I collects every j within the cutoff (FP64 distances, the
// test of the reference: r2 > cutoff^2 is out).  Entries beyond cap are
// counted but not stored; the host grows the list and rebuilds.  For every
// row entry r with J >= I it also writes the lane's tile masks: bit jj for
// each listed j = J * 64 + jj with j > i, in tmaskcc[r * 64 + lane] when i
// and j are both bare charges (chg), in tmask[r * 64 + lane] otherwise.
__global__ __launch_bounds__(64) void k_nl_build(int n, const int* __restrict__ rstart, const int* __restrict__ rows,
                                                 const float4* __restrict__ bmin, const float4* __restrict__ bmax,
                                                 const double4* __restrict__ pos, const int* __restrict__ partners,
                                                 const int* __restrict__ npart, Box<double> box, double rc2, int cap,
                                                 int* __restrict__ nbr, int* __restrict__ count,
                                                 int* __restrict__ maxcount, unsigned long long* __restrict__ tmask,
                                                 int* __restrict__ pnbr, int* __restrict__ pcount,
                                                 int* __restrict__ polcnt, const int* __restrict__ chg,
                                                 unsigned long long* __restrict__ tmaskcc) {
    __shared__ double4 sp[kTile];
    const int I = blockIdx.x, lane = threadIdx.x;
    const int i = I * kTile + lane;
    const bool iv = i < n;
    const int ic = iv ? i : n - 1;
    const double4 pi = pos[ic];
    const int* parts = partners + (size_t)ic * kMaxPartners;
    const int np = iv ? npart[ic] : 0;
    // bounding box of block I (centre, half extent) for the candidate filter
    const float4 b0 = bmin[I], b1 = bmax[I];
    const float cI[3] = {0.5f * (b0.x + b1.x), 0.5f * (b0.y + b1.y), 0.5f * (b0.z + b1.z)};
    const float hI[3] = {0.5f * (b1.x - b0.x), 0.5f * (b1.y - b0.y), 0.5f * (b1.z - b0.z)};
    const float rcf = sqrtf((float)rc2) * 1.0001f + 1e-5f;
    const Box<float> fbox{{(float)box.L[0], (float)box.L[1], (float)box.L[2]},
                          {(float)box.inv[0], (float)box.inv[1], (float)box.inv[2]}};
    const bool ibare = iv && chg[ic] != 0;
    int cnt = 0, pcnt = 0;
    const int r0 = rstart[I], r1 = rstart[I + 1];
    // the next row's atom is loaded one row ahead, so its latency overlaps the current row
    int jnext = r0 < r1 ? min(rows[r0] * kTile + lane, n - 1) : 0;
    double4 pnext = pos[jnext];
    int cnext = chg[jnext];
    for (int r = r0; r < r1; r++) {
        const int J = rows[r];
        const double4 pj = pnext;
        const unsigned long long jbare = __ballot(cnext != 0 && J * kTile + lane < n);
        if (r + 1 < r1) {
            jnext = min(rows[r + 1] * kTile + lane, n - 1);
            pnext = pos[jnext];
            cnext = chg[jnext];
        }
        __syncthreads();
        // each lane stages one atom of J and tests it against block I's box
        // (minimum image of the distance to the box; never rejects a pair
        // within the cutoff, also when the box is wide)
        bool cand;
        {
            const int j = J * kTile + lane;
            sp[lane] = pj;
            const float d[3] = {(float)pj.x - cI[0], (float)pj.y - cI[1], (float)pj.z - cI[2]};
            float d2 = 0.f;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float dk = d[k] - fbox.L[k] * floorf(d[k] * fbox.inv[k] + 0.5f);
                const float g = fmaxf(0.f, fabsf(dk) - hI[k]);
                d2 += g * g;
            }
            cand = j < n && d2 <= rcf * rcf;
        }
        unsigned long long cmask = __ballot(cand);
        __syncthreads();
        unsigned long long skip = 0;
        for (int k = 0; k < np; k++) {
            const int j = parts[k] >> 2;
            if ((j >> 6) == J) skip |= 1ull << (j & 63);
        }
        if (J == I) skip |= 1ull << lane;
        if (!iv) cmask = 0;  // lanes past the last atom list nothing
        // tile-mask bits: j > i, i.e. all of J when J > I, those above the lane when J == I
        const unsigned long long upper = J > I ? ~0ull : (J == I ? ~((2ull << lane) - 1ull) : 0ull);
        unsigned long long tm = 0;
        while (cmask) {
            const int jj = __ffsll((long long)cmask) - 1;
            cmask &= cmask - 1ull;
            if ((skip >> jj) & 1ull) continue;
            const double4 pj = sp[jj];
            const V3<double> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), box);
            if (dot(dr, dr) > rc2) continue;
            if (cnt < cap) nbr[nl_index(i, cnt, n)] = J * kTile + jj;
            cnt++;
            tm |= 1ull << jj;
            if (pj.w != 0.0) {  // j polarisable (damping factor nonzero)
                if (pcnt < cap) pnbr[nl_index(i, pcnt, n)] = J * kTile + jj;
                pcnt++;
            }
        }
        if (J >= I) {
            // pairs of two bare charges go to their own list (their interaction is q q erfc(a r) / r)
            const unsigned long long cc = ibare ? (tm & jbare) : 0ull;
            tmask[(size_t)r * kTile + lane] = (tm & ~cc) & upper;
            tmaskcc[(size_t)r * kTile + lane] = cc & upper;
        }
    }
    if (iv) {
        count[i] = min(cnt, cap);
        pcount[i] = min(pcnt, cap);
        atomicMax(maxcount, cnt);
    }
    const unsigned long long pol = __ballot(iv && pi.w != 0.0);
    if (lane == 0) polcnt[I] = __popcll(pol);
}

// Sorted indices of the polarisable atoms, in order: block I writes its
// polarisable lanes from start[I].
__global__ __launch_bounds__(64) void k_pol_fill(int n, const double4* __restrict__ pos, const int* __restrict__ start,
                                                 int* __restrict__ idx) {
    const int I = blockIdx.x, lane = threadIdx.x;
    const int i = I * kTile + lane;
    const bool p = i < n && pos[i].w != 0.0;
    const unsigned long long m = __ballot(p);
    if (p) idx[start[I] + __popcll(m & ((1ull << lane) - 1ull))] = i;
}

// ---------------------------------------------------------------------------
// list-driven field kernels: kNlSplit threads per atom, each taking every
// kNlSplit-th entry; the partial fields meet in a lane reduction and are
// added once per atom.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void k_nl_fixed_field(NbrList L, const int* __restrict__ atoms, int natoms,
                                                        PairAtoms<R> A, PairParams<R> P, acc_t* __restrict__ field) {
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(g / kNlSplit), sub = (int)(g % kNlSplit);
    const int n = P.n;
    const bool iv = t < natoms;
    const int i = atoms ? atoms[iv ? t : natoms - 1] : (iv ? t : n - 1);
    const int ic = i;
    const R4<R> pi = A.pos[ic];
    const R2<R> tdi = A.td[ic];
    R mi[kMP];
#pragma unroll
    for (int k = 0; k < kMP; k++) mi[k] = A.lab[(size_t)ic * kMP + k];
    const int cnt = iv ? L.count[ic] : 0;
    // entry k of atom i: lp advances by n * kNlSplit per step of kNlSplit
    const int* lp = L.nbr + (size_t)ic * kNlSplit + sub;
    const size_t stride = (size_t)n * kNlSplit;
    int jn = sub < cnt ? lp[0] : 0;
    R ex = 0, ey = 0, ez = 0;
    for (int k = sub; k < cnt; k += kNlSplit) {
        const int j = jn;
        lp += stride;
        if (k + kNlSplit < cnt) jn = *lp;  // next index in flight during this pair
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        R mj[kMP];
#pragma unroll
        for (int c = 0; c < kMP; c++) mj[c] = A.lab[(size_t)j * kMP + c];
        // listed pairs are not covalent partners: scale 1, default Thole width
        const FixedRadial<R> gr = fixed_field_radial<R>(dot(dr, dr), tdi, R2<R>{R(0), pj.w}, R(1), P);
        const V3<R> fi = fixed_field_from<R>(mj, dr, gr);
        ex += fi.x;
        ey += fi.y;
        ez += fi.z;
    }
#pragma unroll
    for (int off = 1; off < kNlSplit; off <<= 1) {
        ex += __shfl_xor(ex, off);
        ey += __shfl_xor(ey, off);
        ez += __shfl_xor(ez, off);
    }
    if (iv && sub == 0) {
        acc_add(&field[3 * (size_t)i + 0], ex);
        acc_add(&field[3 * (size_t)i + 1], ey);
        acc_add(&field[3 * (size_t)i + 2], ez);
    }
}

template <class R>
__global__ __launch_bounds__(256) void k_nl_induced_field(NbrList L, const int* __restrict__ atoms, int natoms,
                                                          PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                                                          acc_t* __restrict__ field) {
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(g / kNlSplit), sub = (int)(g % kNlSplit);
    const int n = P.n;
    const bool iv = t < natoms;
    const int i = atoms ? atoms[iv ? t : natoms - 1] : (iv ? t : n - 1);
    const int ic = i;
    const R4<R> pi = A.pos[ic];
    const int cnt = iv ? L.count[ic] : 0;
    const int* lp = L.nbr + (size_t)ic * kNlSplit + sub;
    const size_t stride = (size_t)n * kNlSplit;
    int jn = sub < cnt ? lp[0] : 0;
    R ex = 0, ey = 0, ez = 0;
    for (int k = sub; k < cnt; k += kNlSplit) {
        const int j = jn;
        lp += stride;
        if (k + kNlSplit < cnt) jn = *lp;
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        const V3<R> uj = mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]);
        // listed pairs are not covalent partners: polarization scale 1, default Thole width
        const V3<R> fi =
            induced_field_from<R>(uj, dr, induced_field_radial<R>(dot(dr, dr), pi.w, pj.w, P.default_thole, P));
        ex += fi.x;
        ey += fi.y;
        ez += fi.z;
    }
#pragma unroll
    for (int off = 1; off < kNlSplit; off <<= 1) {
        ex += __shfl_xor(ex, off);
        ey += __shfl_xor(ey, off);
        ez += __shfl_xor(ez, off);
    }
    if (iv && sub == 0) {
        acc_add(&field[3 * (size_t)i + 0], ex);
        acc_add(&field[3 * (size_t)i + 1], ey);
        acc_add(&field[3 * (size_t)i + 2], ez);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_block_rows(hipStream_t s, int nb, const float4* bmin, const float4* bmax, Box<float> box, float cutoff,
                       const int* start, int* counts, int* rows, int* owner) {
    hipLaunchKernelGGL(k_block_rows, dim3(nb), dim3(kTile), 0, s, nb, bmin, bmax, box, cutoff, start, counts, rows,
                       owner);
}
void launch_exclusive_scan(hipStream_t s, int m, const int* counts, int* start) {
    hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, s, m, counts, start);
}
void launch_nl_build(hipStream_t s, int n, const int* rstart, const int* rows, const float4* bmin, const float4* bmax,
                     const double4* pos, const int* partners, const int* npart, Box<double> box, double rc2, int cap,
                     int* nbr, int* count, int* maxcount, unsigned long long* tmask, int* pnbr, int* pcount,
                     int* polcnt, const int* chg, unsigned long long* tmaskcc) {
    const int nb = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_nl_build, dim3(nb), dim3(kTile), 0, s, n, rstart, rows, bmin, bmax, pos, partners, npart, box,
                       rc2, cap, nbr, count, maxcount, tmask, pnbr, pcount, polcnt, chg, tmaskcc);
}
void launch_pol_fill(hipStream_t s, int n, const double4* pos, const int* start, int* idx) {
    const int nb = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_pol_fill, dim3(nb), dim3(kTile), 0, s, n, pos, start, idx);
}
template <class R>
void launch_nl_fixed_field(hipStream_t s, int natoms, const int* atoms, NbrList l, PairAtoms<R> a, PairParams<R> p,
                           acc_t* field) {
    if (natoms <= 0) return;
    const size_t threads = (size_t)natoms * kNlSplit;
    hipLaunchKernelGGL(k_nl_fixed_field<R>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, l, atoms, natoms,
                       a, p, field);
}
template <class R>
void launch_nl_induced_field(hipStream_t s, int natoms, const int* atoms, NbrList l, PairAtoms<R> a, const R* mu,
                             PairParams<R> p, acc_t* field) {
    if (natoms <= 0) return;
    const size_t threads = (size_t)natoms * kNlSplit;
    hipLaunchKernelGGL(k_nl_induced_field<R>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, l, atoms,
                       natoms, a, mu, p, field);
}

#define MPID_INSTANTIATE(R)                                                                                         \
    template void launch_nl_fixed_field<R>(hipStream_t, int, const int*, NbrList, PairAtoms<R>, PairParams<R>,       \
                                           acc_t*);                                                                \
    template void launch_nl_induced_field<R>(hipStream_t, int, const int*, NbrList, PairAtoms<R>, const R*,          \
                                             PairParams<R>, acc_t*);
MPID_INSTANTIATE(float)
MPID_INSTANTIATE(double)
#undef MPID_INSTANTIATE

}  // namespace mpid
