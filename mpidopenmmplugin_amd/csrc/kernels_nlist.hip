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
//   2. neighbour masks: one wavefront per block pair (I, J >= I) tests each
//      pair once and writes the 64-bit masks of both directions (the second
//      by a bit-matrix transpose), skipping self pairs and scaled covalent
//      partners (those pairs go to the covalent-pair kernels);
//   3. atom lists: lane i of block I walks its row in ascending J and appends
//      the atoms of each mask.
// Both steps are deterministic, so the summation order of every field is.
// The solver needs fields at polarisable sites only: an atom without
// polarisability carries no induced dipole (mu = S y with S = 0) and its
// row of the solver's operator is multiplied by S = 0.  An evaluation
// therefore lists the polarisable atoms only, each with its full list (for
// the permanent field) and a list of its polarisable neighbours (for the
// solver's operator); the inspection entry points, which want fields at
// every atom, ask for full lists of all atoms.
//
// List layout: entry k of list row l sits at ((k / S) * rows + l) * S + k % S
// with S = kNlSplit, so the S threads that share an atom read S adjacent
// words and a wavefront reads 64 consecutive words per step.
#include "pair_math.h"

namespace mpid {

__device__ __forceinline__ size_t nl_index(int l, int k, int rows) {
    return ((size_t)(k / kNlSplit) * rows + l) * kNlSplit + (k % kNlSplit);
}

// Whether an axis-aligned box of half-widths h, centred at separation d
// from a point, comes within rc of that point under some periodic image:
// the point-to-box gap of every image is at most the distance between any
// point of the box and the given point, so a pair within rc passes.  A
// rectangular cell wraps each axis alone.  A triclinic cell (reduced form)
// tries every lattice shift k c + j b + i a that brings the z, then the y,
// then the x separation within rc + h (one shift more on each side for the
// float rounding of the bounds); such a shift exists for any pair within rc,
// whichever image the pair's own minimum image picks.
__device__ __forceinline__ bool gap_within(float dx, float dy, float dz, float hx, float hy, float hz,
                                           const Box<float>& box, float rc) {
    const float rc2 = rc * rc;
    if (box.bx == 0.f && box.cx == 0.f && box.cy == 0.f) {
        const float d[3] = {dx, dy, dz}, h[3] = {hx, hy, hz};
        float d2 = 0.f;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float dk = d[k] - box.L[k] * floorf(d[k] * box.inv[k] + 0.5f);
            const float g = fmaxf(0.f, fabsf(dk) - h[k]);
            d2 += g * g;
        }
        return d2 <= rc2;
    }
    const V3<float> d = min_image(mk(dx, dy, dz), box);
    const int k0 = (int)floorf((-rc - hz - d.z) * box.inv[2]) - 1, k1 = (int)ceilf((rc + hz - d.z) * box.inv[2]) + 1;
    for (int k = k0; k <= k1; k++) {
        const float z = d.z + (float)k * box.L[2];
        const float gz = fmaxf(0.f, fabsf(z) - hz);
        if (gz * gz > rc2) continue;
        const float y0 = d.y + (float)k * box.cy, x0 = d.x + (float)k * box.cx;
        const int j0 = (int)floorf((-rc - hy - y0) * box.inv[1]) - 1, j1 = (int)ceilf((rc + hy - y0) * box.inv[1]) + 1;
        for (int j = j0; j <= j1; j++) {
            const float y = y0 + (float)j * box.L[1];
            const float gy = fmaxf(0.f, fabsf(y) - hy);
            if (gz * gz + gy * gy > rc2) continue;
            const float x1 = x0 + (float)j * box.bx;
            const int i0 = (int)floorf((-rc - hx - x1) * box.inv[0]) - 1,
                      i1 = (int)ceilf((rc + hx - x1) * box.inv[0]) + 1;
            for (int i = i0; i <= i1; i++) {
                const float gx = fmaxf(0.f, fabsf(x1 + (float)i * box.L[0]) - hx);
                if (gz * gz + gy * gy + gx * gx <= rc2) return true;
            }
        }
    }
    return false;
}

// Bounding boxes of blocks I and J within the cutoff (the gap test above on
// the separation of the box centres with the summed half-widths, with a
// little slack for the float rounding of the bounds).
__device__ __forceinline__ bool blocks_within(float4 a0, float4 a1, float4 b0, float4 b1, const Box<float>& box,
                                              float cutoff) {
    const float rc = cutoff * 1.0001f + 1e-5f;
    return gap_within(0.5f * (b0.x + b1.x) - 0.5f * (a0.x + a1.x), 0.5f * (b0.y + b1.y) - 0.5f * (a0.y + a1.y),
                      0.5f * (b0.z + b1.z) - 0.5f * (a0.z + a1.z), 0.5f * (a1.x - a0.x) + 0.5f * (b1.x - b0.x),
                      0.5f * (a1.y - a0.y) + 0.5f * (b1.y - b0.y), 0.5f * (a1.z - a0.z) + 0.5f * (b1.z - b0.z), box,
                      rc);
}

// One wavefront per block I.  Count mode (rows == nullptr) writes counts[I]
// and upper[I], the number of those blocks J with J >= I; fill mode writes
// the row at rows[start[I] ...] in ascending J, and I into owner[] at the
// same positions.
__global__ __launch_bounds__(64) void k_block_rows(int nb, const float4* __restrict__ bmin,
                                                   const float4* __restrict__ bmax, Box<float> box, float cutoff,
                                                   const int* __restrict__ start, int* __restrict__ counts,
                                                   int* __restrict__ upper, int* __restrict__ rows,
                                                   int* __restrict__ owner, int cap) {
    const int I = blockIdx.x, lane = threadIdx.x;
    const float4 a0 = bmin[I], a1 = bmax[I];
    const int base = rows ? start[I] : 0;
    int c = 0, cu = 0;
    for (int J0 = 0; J0 < nb; J0 += kTile) {
        const int J = J0 + lane;
        const bool hit = J < nb && blocks_within(a0, a1, bmin[J], bmax[J], box, cutoff);
        const unsigned long long m = __ballot(hit);
        if (hit && rows) {
            const int at = base + c + __popcll(m & ((1ull << lane) - 1ull));
            if (at < cap) {  // entries beyond the capacity: the host grows it and builds again
                rows[at] = J;
                owner[at] = I;
            }
        }
        c += __popcll(m);
        cu += __popcll(__ballot(hit && J >= I));
    }
    if (lane == 0 && !rows) {
        counts[I] = c;
        upper[I] = cu;
    }
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

// Phase 1 of the atom lists, one wavefront per row entry r = (I, J), J >= I
// (entries with J < I return): lane i of block I tests the atoms of J (FP64
// distances, the test of the reference: r2 > cutoff^2 is out), skipping
// itself and its scaled covalent partners (those pairs go to the
// covalent-pair kernels), and writes the 64-bit neighbour mask of i within J
// to nmask[r * 64 + lane].  For J > I the 64 x 64 bit matrix is transposed
// (one ballot per column) into the masks of J's atoms within I, written at
// the mirrored entry (J, I) of J's row; the block-box test is symmetric, so
// that entry exists (if not, flag[0] is set and the host raises an error).
// Tile masks: bit jj for each listed j = J * 64 + jj with j > i, in
// tmaskcc[r * 64 + lane] when i and j are both bare charges (chg), in
// tmask[r * 64 + lane] otherwise.  The diagonal entry (I, I) writes
// polmask[I], the polarisable atoms of block I, and their number polcnt[I].
__global__ __launch_bounds__(64) void k_nl_masks(int n, const int* __restrict__ rstart, const int* __restrict__ rows,
                                                 const int* __restrict__ owner, const float4* __restrict__ bmin,
                                                 const float4* __restrict__ bmax, const double4* __restrict__ pos,
                                                 const int* __restrict__ partners, const int* __restrict__ npart,
                                                 int pstride, Box<double> box, double rc2,
                                                 const int* __restrict__ chg,
                                                 unsigned long long* __restrict__ nmask,
                                                 unsigned long long* __restrict__ tmask,
                                                 unsigned long long* __restrict__ tmaskcc,
                                                 int* __restrict__ flag, int cap, int4* __restrict__ act,
                                                 const int* __restrict__ ustart) {
    __shared__ double4 sp[kTile];
    const int r = blockIdx.x, lane = threadIdx.x;
    const int nbk = (n + kTile - 1) / kTile;
    if (r >= min(rstart[nbk], cap)) return;  // uniform: the grid may cover the capacity, not the count
    const int I = owner[r], J = rows[r];
    if (J < I) return;  // uniform: written by the mirrored entry
    const int i = I * kTile + lane, j = J * kTile + lane;
    const bool iv = i < n, jv = j < n;
    const int ic = iv ? i : n - 1, jc = jv ? j : n - 1;
    const double4 pi = pos[ic], pj = pos[jc];
    const bool ibare = iv && chg[ic] != 0;
    const unsigned long long jbare = __ballot(jv && chg[jc] != 0);
    sp[lane] = pj;
    // candidates: atoms of J whose distance to block I's box is within the
    // cutoff (minimum image; never rejects a pair within the cutoff)
    bool cand;
    {
        const float4 b0 = bmin[I], b1 = bmax[I];
        const float cI[3] = {0.5f * (b0.x + b1.x), 0.5f * (b0.y + b1.y), 0.5f * (b0.z + b1.z)};
        const float hI[3] = {0.5f * (b1.x - b0.x), 0.5f * (b1.y - b0.y), 0.5f * (b1.z - b0.z)};
        const float rcf = sqrtf((float)rc2) * 1.0001f + 1e-5f;
        Box<float> fb;
        for (int k = 0; k < 3; k++) {
            fb.L[k] = (float)box.L[k];
            fb.inv[k] = (float)box.inv[k];
        }
        fb.bx = (float)box.bx;
        fb.cx = (float)box.cx;
        fb.cy = (float)box.cy;
        cand = jv && gap_within((float)pj.x - cI[0], (float)pj.y - cI[1], (float)pj.z - cI[2], hI[0], hI[1], hI[2],
                                fb, rcf);
    }
    unsigned long long cmask = __ballot(cand);
    unsigned long long skip = J == I ? 1ull << lane : 0ull;
    const int np = iv ? npart[ic] : 0;
    const int* parts = partners + (size_t)ic * pstride;
    for (int k = 0; k < np; k++) {
        const int p = parts[k] >> 2;
        if ((p >> 6) == J) skip |= 1ull << (p & 63);
    }
    if (!iv) cmask = 0;  // lanes past the last atom list nothing
    __syncthreads();
    unsigned long long m = 0;
    while (cmask) {
        const int jj = __ffsll((long long)cmask) - 1;
        cmask &= cmask - 1ull;
        const double4 q = sp[jj];
        const V3<double> dr = min_image(mk(q.x - pi.x, q.y - pi.y, q.z - pi.z), box);
        if (dot(dr, dr) <= rc2) m |= 1ull << jj;
    }
    m &= ~skip;
    nmask[(size_t)r * kTile + lane] = m;
    // tile-mask bits: j > i, i.e. all of J when J > I, those above the lane when J == I;
    // pairs of two bare charges go to their own list (their interaction is q q erfc(a r) / r)
    const unsigned long long upper = J > I ? ~0ull : ~((2ull << lane) - 1ull);
    const unsigned long long cc = ibare ? (m & jbare) : 0ull;
    tmask[(size_t)r * kTile + lane] = (m & ~cc) & upper;
    tmaskcc[(size_t)r * kTile + lane] = cc & upper;
    // the tile's entry in the electrostatics kernels' work list: the row
    // entries J >= I of row I are its last upper[I] entries, placed at
    // ustart[I] (the exclusive scan of upper) in the same order; a tile
    // without pairs is marked r = -1
    const bool has_pairs = __ballot((m & upper) != 0ull) != 0ull;
    if (lane == 0) act[ustart[I + 1] - (rstart[I + 1] - r)] = make_int4(has_pairs ? r : -1, I, J, 0);
    if (J > I) {
        unsigned long long t = 0;
        for (int k = 0; k < kTile; k++) {
            const unsigned long long col = __ballot((m >> k) & 1ull);
            if (lane == k) t = col;
        }
        // the mirrored entry: I in the ascending row of J
        int lo = min(rstart[J], cap), hi = min(rstart[J + 1], cap) - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (rows[mid] < I) lo = mid + 1;
            else hi = mid;
        }
        if (lo <= hi && rows[lo] == I) nmask[(size_t)lo * kTile + lane] = t;
        else if (lane == 0) atomicOr(flag, 1);
    }
}

// Phase 2, one workgroup of kFillWaves wavefronts per block I: the row of I
// is cut into kFillWaves contiguous chunks in ascending J.  Each wavefront
// counts its chunk's entries per atom, and a prefix over the chunks gives
// each its offsets.  The entries are then written atom by atom: for each
// row entry, the wavefront takes in turn each stored atom whose mask is not
// empty, and lane jj writes neighbour jj at the atom's offset plus the
// number of mask bits below jj (all lanes work; a lane-per-atom walk over
// the mask bits would idle the lanes of non-stored atoms and of short
// masks).  all != 0: the full list of every atom, stored by atom index.
// all == 0: for the polarisable atoms only, the full list and the list of
// polarisable neighbours (polmask of the neighbour's block), stored by the
// atom's position p in the sorted index of polarisable atoms
// (p = polstart[I] + rank in the block, npol = polstart[nb] rows).
// count[i] (and pcount[i] when all == 0) are stored for every atom.  Entries
// beyond cap are counted but not stored; maxcount takes the largest stored
// list's length, and the host grows the lists and rebuilds.
constexpr int kFillWaves = 16;
__global__ __launch_bounds__(kFillWaves * kTile) void k_nl_fill(
    int n, int all, const int* __restrict__ rstart, const int* __restrict__ rows,
    const unsigned long long* __restrict__ polmask, const int* __restrict__ polstart,
    const unsigned long long* __restrict__ nmask, int cap, int* __restrict__ nbr, int* __restrict__ count,
    int* __restrict__ maxcount, int* __restrict__ pnbr, int* __restrict__ pcount, int rcap) {
    __shared__ int scnt[kFillWaves][kTile], spcnt[kFillWaves][kTile];
    const int I = blockIdx.x, lane = threadIdx.x & (kTile - 1), w = threadIdx.x / kTile;
    const int nb = (n + kTile - 1) / kTile;
    const int i = I * kTile + lane;
    const bool iv = i < n;
    const unsigned long long pmI = polmask[I];
    const bool ipol = ((pmI >> lane) & 1ull) != 0;
    const bool store = iv && (all || ipol);
    const int li = all ? i : polstart[I] + __popcll(pmI & ((1ull << lane) - 1ull));
    const int nrows = all ? n : polstart[nb];
    // (rows beyond the capacity rcap were not stored; the host grows it and builds again)
    const int r0 = min(rstart[I], rcap), r1 = min(rstart[I + 1], rcap);
    const int per = (r1 - r0 + kFillWaves - 1) / kFillWaves;
    const int a = min(r0 + w * per, r1), b = min(a + per, r1);
    int c = 0, pc = 0;
    for (int r = a; r < b; r++) {
        const unsigned long long m = iv ? nmask[(size_t)r * kTile + lane] : 0ull;
        c += __popcll(m);
        pc += __popcll(m & polmask[rows[r]]);
    }
    scnt[w][lane] = c;
    spcnt[w][lane] = pc;
    __syncthreads();
    int cnt = 0, pcnt = 0, tot = 0, ptot = 0;
    for (int v = 0; v < kFillWaves; v++) {
        if (v == w) {
            cnt = tot;
            pcnt = ptot;
        }
        tot += scnt[v][lane];
        ptot += spcnt[v][lane];
    }
    const unsigned long long below = (1ull << lane) - 1ull;  // mask bits of the lanes below this one
    for (int r = a; r < b; r++) {
        const unsigned long long mrow = iv ? nmask[(size_t)r * kTile + lane] : 0ull;
        const int J = rows[r];
        const unsigned long long polJ = all ? 0ull : polmask[J];
        unsigned long long act = __ballot(store && mrow != 0ull);
        while (act) {
            const int t = __ffsll((long long)act) - 1;  // uniform: the atom of lane t
            act &= act - 1ull;
            const unsigned lo = __builtin_amdgcn_readlane((unsigned)mrow, t);
            const unsigned hi = __builtin_amdgcn_readlane((unsigned)(mrow >> 32), t);
            const unsigned long long m = ((unsigned long long)hi << 32) | lo;
            const int base = __builtin_amdgcn_readlane(cnt, t);
            const int lt = __builtin_amdgcn_readlane(li, t);
            if ((m >> lane) & 1ull) {
                const int k = base + __popcll(m & below);
                if (k < cap) nbr[nl_index(lt, k, nrows)] = J * kTile + lane;
            }
            const unsigned long long mp = m & polJ;
            if (!all) {
                const int pbase = __builtin_amdgcn_readlane(pcnt, t);
                if ((mp >> lane) & 1ull) {
                    const int k = pbase + __popcll(mp & below);
                    if (k < cap) pnbr[nl_index(lt, k, nrows)] = J * kTile + lane;
                }
            }
            if (lane == t) {
                cnt += __popcll(m);
                pcnt += __popcll(mp);
            }
        }
    }
    if (w == 0) {
        if (iv) {
            count[i] = min(tot, cap);
            if (!all) pcount[i] = min(ptot, cap);
        }
        // the block's longest list, then one atomic per block (atomics from
        // every lane on one word serialise across the chip)
        int lmax = store ? tot : 0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lmax = max(lmax, __shfl_xor(lmax, off));
        if (lane == 0 && lmax > 0) atomicMax(maxcount, lmax);
    }
}

// The polarisable atoms of each block (damping factor nonzero): a mask and
// its count, one wavefront per block.
__global__ __launch_bounds__(64) void k_pol_mask(int n, const double4* __restrict__ pos,
                                                 unsigned long long* __restrict__ polmask, int* __restrict__ polcnt) {
    const int I = blockIdx.x, lane = threadIdx.x;
    const int i = I * kTile + lane;
    const unsigned long long pm = __ballot(i < n && pos[i < n ? i : n - 1].w != 0.0);
    if (lane == 0) {
        polmask[I] = pm;
        polcnt[I] = __popcll(pm);
    }
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
    // lists of an atoms index are stored by position in the index
    const int li = atoms ? (iv ? t : natoms - 1) : ic;
    const int* lp = L.nbr + (size_t)li * kNlSplit + sub;
    const size_t stride = (size_t)(atoms ? natoms : n) * kNlSplit;
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

// GRAD: also the field gradient (xx yy zz xy xz yz) into grad[6 i ..]
// (Extrapolated polarization).
template <class R, bool GRAD>
__global__ __launch_bounds__(256) void k_nl_induced_field(NbrList L, const int* __restrict__ atoms, int natoms,
                                                          PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                                                          acc_t* __restrict__ field, acc_t* __restrict__ grad) {
    if (skipped(P.skip)) return;
    const unsigned g = (unsigned)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int t = (int)(g / kNlSplit), sub = (int)(g % kNlSplit);
    const int n = P.n;
    const bool iv = t < natoms;
    const int i = atoms ? atoms[iv ? t : natoms - 1] : (iv ? t : n - 1);
    const int ic = i;
    const R4<R> pi = A.pos[ic];
    const int cnt = iv ? L.count[ic] : 0;
    // lists of an atoms index are stored by position in the index
    const int li = atoms ? (iv ? t : natoms - 1) : ic;
    const int* lp = L.nbr + (size_t)li * kNlSplit + sub;
    const size_t stride = (size_t)(atoms ? natoms : n) * kNlSplit;
    int jn = sub < cnt ? lp[0] : 0;
    R ex = 0, ey = 0, ez = 0;
    R gs[6] = {0, 0, 0, 0, 0, 0};
    for (int k = sub; k < cnt; k += kNlSplit) {
        const int j = jn;
        lp += stride;
        if (k + kNlSplit < cnt) jn = *lp;
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        const V3<R> uj = mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]);
        // listed pairs are not covalent partners: polarization scale 1, default Thole width
        V3<R> fi;
        if constexpr (GRAD) {
            const InducedRadialGrad<R> f = induced_grad_radial<R>(dot(dr, dr), pi.w, pj.w, P.default_thole, P);
            fi = induced_field_from<R>(uj, dr, InducedRadial<R>{f.f1, f.f2});
            induced_grad_add<R>(uj, dr, f.f2, f.f3, gs);
        } else {
            fi = induced_field_from<R>(uj, dr, induced_field_radial<R>(dot(dr, dr), pi.w, pj.w, P.default_thole, P));
        }
        ex += fi.x;
        ey += fi.y;
        ez += fi.z;
    }
#pragma unroll
    for (int off = 1; off < kNlSplit; off <<= 1) {
        ex += __shfl_xor(ex, off);
        ey += __shfl_xor(ey, off);
        ez += __shfl_xor(ez, off);
        if constexpr (GRAD) {
#pragma unroll
            for (int c = 0; c < 6; c++) gs[c] += __shfl_xor(gs[c], off);
        }
    }
    if (iv && sub == 0) {
        acc_add(&field[3 * (size_t)i + 0], ex);
        acc_add(&field[3 * (size_t)i + 1], ey);
        acc_add(&field[3 * (size_t)i + 2], ez);
        if constexpr (GRAD) {
#pragma unroll
            for (int c = 0; c < 6; c++) acc_add(&grad[6 * (size_t)i + c], gs[c]);
        }
    }
}

// Thole damping derivatives of the energy gradients: h = u dE/du of each
// listed pair of polarisable sites (thole_u_derivative; listed pairs are not
// covalent partners: polarization scale 1, default Thole width), summed over
// the atom's polarisable neighbours into hd[i] and hdef[i].  Each pair is
// visited from both of its atoms.
template <class R>
__global__ __launch_bounds__(256) void k_nl_thole(NbrList L, const int* __restrict__ atoms, int natoms, PairAtoms<R> A,
                                                  const R* __restrict__ mu, PairParams<R> P, acc_t* __restrict__ hd,
                                                  acc_t* __restrict__ hdef) {
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(g / kNlSplit), sub = (int)(g % kNlSplit);
    const bool iv = t < natoms;
    const int i = atoms[iv ? t : natoms - 1];
    const R4<R> pi = A.pos[i];
    R mi[kMP];
#pragma unroll
    for (int k = 0; k < kMP; k++) mi[k] = A.lab[(size_t)i * kMP + k];
    const V3<R> ui = mk(mu[3 * (size_t)i], mu[3 * (size_t)i + 1], mu[3 * (size_t)i + 2]);
    const int cnt = iv ? L.count[i] : 0;
    const int* lp = L.nbr + (size_t)(iv ? t : natoms - 1) * kNlSplit + sub;
    const size_t stride = (size_t)natoms * kNlSplit;
    int jn = sub < cnt ? lp[0] : 0;
    R h = 0;
    for (int k = sub; k < cnt; k += kNlSplit) {
        const int j = jn;
        lp += stride;
        if (k + kNlSplit < cnt) jn = *lp;
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        R mj[kMP];
#pragma unroll
        for (int c = 0; c < kMP; c++) mj[c] = A.lab[(size_t)j * kMP + c];
        const V3<R> uj = mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]);
        h += thole_u_derivative<R>(dr, dot(dr, dr), mi, mj, ui, uj, pi.w, pj.w, P.default_thole, R(1), P.mutual != 0);
    }
#pragma unroll
    for (int off = 1; off < kNlSplit; off <<= 1) h += __shfl_xor(h, off);
    if (iv && sub == 0 && h != R(0)) {
        acc_add(&hd[i], h);
        acc_add(&hdef[i], h);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_block_rows(hipStream_t s, int nb, const float4* bmin, const float4* bmax, Box<float> box, float cutoff,
                       const int* start, int* counts, int* upper, int* rows, int* owner, int cap) {
    hipLaunchKernelGGL(k_block_rows, dim3(nb), dim3(kTile), 0, s, nb, bmin, bmax, box, cutoff, start, counts, upper,
                       rows, owner, cap);
}
void launch_exclusive_scan(hipStream_t s, int m, const int* counts, int* start) {
    hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, s, m, counts, start);
}
void launch_nl_build(hipStream_t s, int n, int nrows, int rcap, bool all, const int* rstart, const int* rows,
                     const int* owner, const float4* bmin, const float4* bmax, const double4* pos,
                     const int* partners, const int* npart, int pstride, Box<double> box, double rc2, int cap,
                     unsigned long long* nmask, unsigned long long* polmask, int* polcnt, int* polstart, int* nbr,
                     int* count, int* scalars, unsigned long long* tmask, int* pnbr, int* pcount, const int* chg,
                     unsigned long long* tmaskcc, int4* act, const int* ustart) {
    const int nb = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_pol_mask, dim3(nb), dim3(kTile), 0, s, n, pos, polmask, polcnt);
    if (nrows > 0)
        hipLaunchKernelGGL(k_nl_masks, dim3(nrows), dim3(kTile), 0, s, n, rstart, rows, owner, bmin, bmax, pos,
                           partners, npart, pstride, box, rc2, chg, nmask, tmask, tmaskcc, scalars + 1, rcap, act,
                           ustart);
    launch_exclusive_scan(s, nb, polcnt, polstart);
    hipLaunchKernelGGL(k_nl_fill, dim3(nb), dim3(kFillWaves * kTile), 0, s, n, all ? 1 : 0, rstart, rows, polmask,
                       polstart, nmask, cap, nbr, count, scalars, pnbr, pcount, rcap);
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
                             PairParams<R> p, acc_t* field, acc_t* grad) {
    if (natoms <= 0) return;
    const size_t threads = (size_t)natoms * kNlSplit;
    const dim3 grid((unsigned)((threads + 255) / 256));
    if (grad)
        hipLaunchKernelGGL((k_nl_induced_field<R, true>), grid, dim3(256), 0, s, l, atoms, natoms, a, mu, p, field,
                           grad);
    else
        hipLaunchKernelGGL((k_nl_induced_field<R, false>), grid, dim3(256), 0, s, l, atoms, natoms, a, mu, p, field,
                           nullptr);
}

template <class R>
void launch_nl_thole(hipStream_t s, int natoms, const int* atoms, NbrList l, PairAtoms<R> a, const R* mu,
                     PairParams<R> p, acc_t* hd, acc_t* hdef) {
    if (natoms <= 0) return;
    const size_t threads = (size_t)natoms * kNlSplit;
    hipLaunchKernelGGL(k_nl_thole<R>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, l, atoms, natoms, a,
                       mu, p, hd, hdef);
}

#define MPID_INSTANTIATE(R)                                                                                         \
    template void launch_nl_thole<R>(hipStream_t, int, const int*, NbrList, PairAtoms<R>, const R*, PairParams<R>,  \
                                     acc_t*, acc_t*);                                                              \
    template void launch_nl_fixed_field<R>(hipStream_t, int, const int*, NbrList, PairAtoms<R>, PairParams<R>,       \
                                           acc_t*);                                                                \
    template void launch_nl_induced_field<R>(hipStream_t, int, const int*, NbrList, PairAtoms<R>, const R*,          \
                                             PairParams<R>, acc_t*, acc_t*);
MPID_INSTANTIATE(float)
MPID_INSTANTIATE(double)
#undef MPID_INSTANTIATE

}  // namespace mpid
