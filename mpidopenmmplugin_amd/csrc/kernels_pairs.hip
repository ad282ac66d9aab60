// Real-space electrostatics (energy, forces, torques) of the MPIDForce PME
// path on gfx950, and the covalent-pair kernels.
//
// Atoms are sorted along a Morton curve and cut into blocks of 64 (one
// wavefront).  A tile is a pair of blocks (I, J >= I) whose bounding boxes
// come within the cutoff; the neighbour-list build leaves, for every tile and
// every atom i of block I, a 64-bit mask of the atoms j of block J that are
// within the cutoff and are not covalent partners of i (j > i when I == J).
// One workgroup (four wavefronts) owns one tile: it ranks the set bits with
// prefix sums and writes the tile's pairs into an LDS list, sorted into three
// classes (two bare charges, one bare charge, general), and then walks each
// class in chunks of 64 entries, one chunk per wavefront.  Every thread thus
// works on a pair inside the cutoff, whatever the tile's fill, and each class
// runs its own formula without divergence.  Both sides accumulate in LDS
// (atomic adds) and are flushed to the fixed-point global accumulators once
// per tile.
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

constexpr int kElecThreads = 256;  // four wavefronts share one tile

// kPot (the energy gradients): also the potential vectors dE/dm of every
// pair (20 stored lab components per atom, induced dipoles fixed) into pot;
// all_general: every pair takes the general formula, so that sites that are
// bare charges get all 20 components (else they get the charge one only).
template <class R, bool kPot>
__device__ __forceinline__ void electrostatics_tile(const int4* __restrict__ act, const int* __restrict__ nact,
                                                                  const int* __restrict__ rows,
                                                                  const int* __restrict__ owner,
                                                                  const unsigned long long* __restrict__ tmask,
                                                                  const unsigned long long* __restrict__ tmaskcc,
                                                                  const int* __restrict__ chg,
                                                                  PairAtoms<R> A, const R* __restrict__ mu,
                                                                  PairParams<R> P, const double4* __restrict__ pos64,
                                                                  const double* __restrict__ lab64,
                                                                  PairParams<double> P64, acc_t* __restrict__ force,
                                                                  acc_t* __restrict__ torque,
                                                                  acc_t* __restrict__ energy, acc_t* __restrict__ pot,
                                                                  int all_general) {
    // Multipole rows of 20 reals, read as 16-byte vectors: with a 20-word
    // stride, 16 lanes reading 16 different rows cover each bank once.
    constexpr int kS = kMP;
    constexpr int kV = 16 / (int)sizeof(R);  // reals per 16-byte vector
    // the tile's pairs as (il << 8 | jj), at most 64 x 64 in all
    __shared__ unsigned short splist[kTile * kTile];
    // [side][component][atom], side 0: atoms of block I, 1: atoms of block J;
    // components: force x, y, z, torque x, y, z.  FP64 in both precisions
    // (LDS atomic adds of doubles keep up on gfx950, those of floats stall).
    __shared__ double acc[2][6][kTile];
    // potential vectors (kPot): [side][component][atom]
    __shared__ double pacc[kPot ? 2 : 1][kPot ? kMP : 1][kTile];
    __shared__ R4<R> spos[2][kTile];
    __shared__ __align__(16) R smp[2][kTile * kS];
    __shared__ R smu[2][kTile * 3];
    __shared__ int san[2][kTile];
    __shared__ double4 spos64[2][kTile];  // x, y, z, charge in FP64 for the charge-charge channel
    __shared__ double sen[kElecThreads / kTile];
    const int tid = threadIdx.x, lane = tid & (kTile - 1), wave = tid / kTile;
    const int n = P.n;
    // One workgroup per entry {r, I, J} of the list of block pairs (J >= I);
    // the launch covers an upper bound of its length (the entries past it
    // lie inside the buffer and are not used), and r = -1 marks a block
    // pair without pairs.
    const int w = blockIdx.x;
    const int4 tile = act[w];
    if (w >= *nact || tile.x < 0) return;
    const int r = tile.x, I = tile.y, J = tile.z;
    // A workgroup's wavefronts sit on fixed SIMDs of a CU, and a tile's work
    // is uneven: one wavefront takes the general pairs (about four times the
    // cost of a pair with a bare charge).  The roles therefore rotate over
    // the wavefronts from tile to tile (a multiplicative hash of the tile
    // index), so that the workgroups sharing a CU load its SIMDs alike.
    const int role = (wave - (int)(((unsigned)r * 0x9E3779B1u) >> 30)) & 3;

    // Every load of the tile is issued here, before any of it is used, so
    // that their latencies overlap: the masks and bare-charge flags for the
    // ranking; roles 0 and 3 the per-atom fields of sides 0 and 1; and the
    // multipole rows, two threads per atom (thread t: atom t mod 64, side
    // (t / 64) mod 2, half t / 128 of the row's 16-byte vectors).  A bare
    // charge's row is its charge and zeros (its other components are zero),
    // so only its first vector is read; atoms past the last read as zeros.
    bool bi = chg[min(I * kTile + lane, n - 1)] != 0;
    const bool bjl = chg[min(J * kTile + lane, n - 1)] != 0;
    unsigned long long mc = tmaskcc[(size_t)r * kTile + lane];
    unsigned long long mg = tmask[(size_t)r * kTile + lane];
    const bool stager = role == 0 || role == 3;
    const int sside = role == 0 ? 0 : 1;
    const int sa = min((sside == 0 ? I : J) * kTile + lane, n - 1);
    R4<R> st_pos{};
    R st_mu[3] = {};
    int st_an = 0;
    double4 st_p64{};
    double st_q = 0.0;
    if (stager) {
        st_pos = A.pos[sa];
        st_mu[0] = mu[3 * (size_t)sa + 0];
        st_mu[1] = mu[3 * (size_t)sa + 1];
        st_mu[2] = mu[3 * (size_t)sa + 2];
        st_an = A.aniso[sa];
        st_p64 = pos64[sa];
        st_q = lab64[(size_t)sa * kMP];
    }
    typedef R vec __attribute__((ext_vector_type(kV)));
    static_assert(kElecThreads == 4 * kTile, "two sides x two halves of the multipole rows");
    constexpr int kVA = kMP / kV;            // vectors per row
    constexpr int kHalf = (kVA + 1) / 2;     // vectors per thread
    const int cside = (tid >> 6) & 1, cpart = tid >> 7;
    const int catom = (cside == 0 ? I : J) * kTile + lane;
    const bool cbare = cside == 0 ? bi : bjl;
    vec st_mp[kHalf];
    {
        const vec* row = reinterpret_cast<const vec*>(A.lab + (size_t)min(catom, n - 1) * kMP);
#pragma unroll
        for (int u = 0; u < kHalf; u++) {
            const int v = cpart * kHalf + u;
            st_mp[u] = (catom < n && v < kVA && (v == 0 || !cbare)) ? row[v] : vec(R(0));
        }
    }

    // Every wavefront ranks the tile's pairs (prefix sums of the lanes'
    // popcounts): pairs of two bare charges at [0, tc), general pairs at
    // [tc, tc + t2), pairs with one bare charge after.  Bit 15 of a one-bare
    // entry says that i (not j) is the bare site.  Roles 1, 2 and 3 write out
    // one class each.
    unsigned long long bj = __ballot(bjl);
    if (kPot && all_general) {
        mg |= mc;
        mc = 0ull;
        bi = false;
        bj = 0ull;
    }
    const unsigned long long m1 = bi ? mg : (mg & bj);
    const unsigned long long m2 = bi ? 0ull : (mg & ~bj);
    const int cc = __popcll(mc), c1 = __popcll(m1), c2 = __popcll(m2);
    int ic = cc, i1 = c1, i2 = c2;
#pragma unroll
    for (int off = 1; off < kTile; off <<= 1) {
        const int vc = __shfl_up(ic, off), v1 = __shfl_up(i1, off), v2 = __shfl_up(i2, off);
        if (lane >= off) {
            ic += vc;
            i1 += v1;
            i2 += v2;
        }
    }
    const int tc = __shfl(ic, kTile - 1), t1 = __shfl(i1, kTile - 1), t2 = __shfl(i2, kTile - 1);
    if (tc + t1 + t2 == 0) return;  // uniform over the workgroup
    if (role == 1) {
        int o = ic - cc;
        for (unsigned long long b = mc; b; b &= b - 1ull)
            splist[o++] = (unsigned short)((lane << 8) | (__ffsll((long long)b) - 1));
    } else if (role == 2) {
        int o = tc + t2 + i1 - c1;
        const int flag = bi ? 0x8000 : 0;
        for (unsigned long long b = m1; b; b &= b - 1ull)
            splist[o++] = (unsigned short)(flag | (lane << 8) | (__ffsll((long long)b) - 1));
    } else if (role == 3) {
        int o = tc + i2 - c2;
        for (unsigned long long b = m2; b; b &= b - 1ull)
            splist[o++] = (unsigned short)((lane << 8) | (__ffsll((long long)b) - 1));
    }
    // the staged loads into LDS, and the accumulators cleared
    if (stager) {
        spos[sside][lane] = st_pos;
        smu[sside][3 * lane + 0] = st_mu[0];
        smu[sside][3 * lane + 1] = st_mu[1];
        smu[sside][3 * lane + 2] = st_mu[2];
        san[sside][lane] = st_an;
        spos64[sside][lane] = make_double4(st_p64.x, st_p64.y, st_p64.z, st_q);
#pragma unroll
        for (int k = 0; k < 6; k++) acc[sside][k][lane] = 0.0;
        if constexpr (kPot) {
#pragma unroll
            for (int k = 0; k < kMP; k++) pacc[sside][k][lane] = 0.0;
        }
    }
#pragma unroll
    for (int u = 0; u < kHalf; u++) {
        const int v = cpart * kHalf + u;
        if (v < kVA) reinterpret_cast<vec*>(&smp[cside][0])[lane * kVA + v] = st_mp[u];
    }
    __syncthreads();
    // Within a wavefront's chunk of 64 entries, lane l takes entry
    // 8 (l mod 8) + l / 8: entries of one atom i are adjacent in the list,
    // and spreading them over the lanes keeps the atomic adds of one
    // instruction off a common address.
    const int perm = ((lane & 7) << 3) | (lane >> 3);
    double en = 0.0;
    // The charge-charge channel runs in FP64 in both precision modes for
    // every pair: it carries most of the magnitude of the real-space energy,
    // and FP32 rounding of the charges, alpha and the Coulomb constant would
    // bias all its terms alike.  This loop takes it for the pairs of two bare
    // charges (where it is the whole interaction) and for the general pairs
    // (apart from their formula, whose registers it would otherwise hold);
    // the one-bare loop takes it with the rest.  Chunks start at role 3.
    for (int q = ((3 - role) & 3) * kTile + perm; q < tc + t2; q += kElecThreads) {
        const int ent = splist[q], il = ent >> 8, jj = ent & 0xff;
        const double4 pi = spos64[0][il], pj = spos64[1][jj];
        const V3<double> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P64.box);
        V3<double> f;
        double e;
        if constexpr (kPot) {
            double phi_i = 0.0, phi_j = 0.0;
            charge_pair<double, true>(dr, dot(dr, dr), pi.w, pj.w, P64, f, e, &phi_i, &phi_j);
            atomicAdd(&pacc[0][0][il], phi_i);
            atomicAdd(&pacc[1][0][jj], phi_j);
        } else {
            charge_pair<double>(dr, dot(dr, dr), pi.w, pj.w, P64, f, e);
        }
        en += e;
        atomicAdd(&acc[0][0][il], -f.x);
        atomicAdd(&acc[0][1][il], -f.y);
        atomicAdd(&acc[0][2][il], -f.z);
        atomicAdd(&acc[1][0][jj], f.x);
        atomicAdd(&acc[1][1][jj], f.y);
        atomicAdd(&acc[1][2][jj], f.z);
    }
    // Pairs with one bare charge b: the FP64 charge-charge channel, and b's
    // charge against a's moments and induced dipole in the working
    // precision (the Cartesian form, elec_pair_bare_cart).  Chunks start at
    // role 1.
    for (int q = ((role - 1) & 3) * kTile + perm; q < t1; q += kElecThreads) {
        const int ent = splist[tc + t2 + q], il = (ent >> 8) & 0x3f, jj = ent & 0xff;
        const bool ibare = (ent & 0x8000) != 0;
        const int sa = ibare ? 1 : 0, ia = ibare ? jj : il;
        const int sb = 1 - sa, ib = ibare ? il : jj;
        const R4<R> pa = spos[sa][ia], pb = spos[sb][ib];
        const V3<R> dr = min_image(mk(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z), P.box);
        R ma[kMP];
        {
            typedef R vec __attribute__((ext_vector_type(kV)));
            const vec* ra = reinterpret_cast<const vec*>(&smp[sa][ia * kS]);
#pragma unroll
            for (int v = 0; v < kMP / kV; v++) {
                const vec x = ra[v];
#pragma unroll
                for (int c = 0; c < kV; c++) ma[v * kV + c] = x[c];
            }
        }
        const V3<R> ua = mk(smu[sa][3 * ia], smu[sa][3 * ia + 1], smu[sa][3 * ia + 2]);
        R pot_a[kPot ? kMP : 1] = {};
        R pot_b0 = R(0);
        const ElecBareOut<R> o = elec_pair_bare_cart<R, kPot>(dr, dot(dr, dr), ma, smp[sb][ib * kS], ua,
                                                              san[sa][ia] != 0, P, pot_a, &pot_b0);
        const double4 qa = spos64[sa][ia], qb = spos64[sb][ib];
        const V3<double> dr64 = min_image(mk(qb.x - qa.x, qb.y - qa.y, qb.z - qa.z), P64.box);
        V3<double> f;
        double e;
        if constexpr (kPot) {
            double phi_a = 0.0, phi_b = 0.0;
            charge_pair<double, true>(dr64, dot(dr64, dr64), qa.w, qb.w, P64, f, e, &phi_a, &phi_b);
            atomicAdd(&pacc[sa][0][ia], phi_a);
            atomicAdd(&pacc[sb][0][ib], phi_b + (double)pot_b0);
#pragma unroll
            for (int k = 1; k < kMP; k++) atomicAdd(&pacc[sa][k][ia], (double)pot_a[k]);
        } else {
            charge_pair<double>(dr64, dot(dr64, dr64), qa.w, qb.w, P64, f, e);
        }
        en += e + (double)o.energy;
        f.x += (double)o.f.x;
        f.y += (double)o.f.y;
        f.z += (double)o.f.z;
        atomicAdd(&acc[sa][0][ia], -f.x);
        atomicAdd(&acc[sa][1][ia], -f.y);
        atomicAdd(&acc[sa][2][ia], -f.z);
        atomicAdd(&acc[sa][3][ia], (double)o.ta.x);
        atomicAdd(&acc[sa][4][ia], (double)o.ta.y);
        atomicAdd(&acc[sa][5][ia], (double)o.ta.z);
        atomicAdd(&acc[sb][0][ib], f.x);
        atomicAdd(&acc[sb][1][ib], f.y);
        atomicAdd(&acc[sb][2][ib], f.z);
    }
    // General pairs: the quasi-internal-frame interaction without its
    // charge-charge channel, in the working precision.  Chunks start at role 0.
    for (int q = role * kTile + perm; q < t2; q += kElecThreads) {
        const int ent = splist[tc + q], il = ent >> 8, jj = ent & 0xff;
        const R4<R> pi = spos[0][il], pj = spos[1][jj];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        R mi[kMP], mj[kMP];
        {
            typedef R vec __attribute__((ext_vector_type(kV)));
            const vec* ri = reinterpret_cast<const vec*>(&smp[0][il * kS]);
            const vec* rj = reinterpret_cast<const vec*>(&smp[1][jj * kS]);
#pragma unroll
            for (int v = 0; v < kMP / kV; v++) {
                const vec a = ri[v], b = rj[v];
#pragma unroll
                for (int c = 0; c < kV; c++) {
                    mi[v * kV + c] = a[c];
                    mj[v * kV + c] = b[c];
                }
            }
        }
        const V3<R> ui = mk(smu[0][3 * il], smu[0][3 * il + 1], smu[0][3 * il + 2]);
        const V3<R> uj = mk(smu[1][3 * jj], smu[1][3 * jj + 1], smu[1][3 * jj + 2]);
        R pot_i[kPot ? kMP : 1] = {}, pot_j[kPot ? kMP : 1] = {};
        const ElecOut<R> o = elec_pair<R, kPot>(dr, dot(dr, dr), mi, mj, ui, uj, R2<R>{R(0), pi.w},
                                                R2<R>{R(0), pj.w}, san[0][il] != 0, san[1][jj] != 0, R(1), R(1), P,
                                                false, pot_i, pot_j);
        if constexpr (kPot) {
#pragma unroll
            for (int k = 0; k < kMP; k++) {
                atomicAdd(&pacc[0][k][il], (double)pot_i[k]);
                atomicAdd(&pacc[1][k][jj], (double)pot_j[k]);
            }
        }
        en += (double)o.energy;
        const V3<R> f = o.f;
        atomicAdd(&acc[0][0][il], -(double)f.x);
        atomicAdd(&acc[0][1][il], -(double)f.y);
        atomicAdd(&acc[0][2][il], -(double)f.z);
        atomicAdd(&acc[0][3][il], (double)o.ti.x);
        atomicAdd(&acc[0][4][il], (double)o.ti.y);
        atomicAdd(&acc[0][5][il], (double)o.ti.z);
        atomicAdd(&acc[1][0][jj], (double)f.x);
        atomicAdd(&acc[1][1][jj], (double)f.y);
        atomicAdd(&acc[1][2][jj], (double)f.z);
        atomicAdd(&acc[1][3][jj], (double)o.tj.x);
        atomicAdd(&acc[1][4][jj], (double)o.tj.y);
        atomicAdd(&acc[1][5][jj], (double)o.tj.z);
    }
    // wavefront sums of the pair energies; every wavefront's pair adds are in LDS
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) en += __shfl_xor(en, off);
    if (lane == 0) sen[wave] = en;
    __syncthreads();
    // flush: role 3 side 0, role 1 side 1 (the two lightest pair loads)
    if (role == 3 || role == 1) {
        const int side = role == 3 ? 0 : 1;
        const int a = (side == 0 ? I : J) * kTile + lane;
        if (a < n) {
            // component by component: a bare charge takes no torque, and an
            // atom without pairs in the tile nothing at all
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double f = acc[side][k][lane], t = acc[side][3 + k][lane];
                if (f != 0.0) acc_add(&force[3 * (size_t)a + k], f);
                if (t != 0.0) acc_add(&torque[3 * (size_t)a + k], t);
            }
            if constexpr (kPot) {
#pragma unroll
                for (int k = 0; k < kMP; k++) {
                    const double v = pacc[side][k][lane];
                    if (v != 0.0) acc_add(&pot[kMP * (size_t)a + k], v);
                }
            }
        }
    }
    if (tid == 0) {
        double e = 0.0;
        for (int k = 0; k < kElecThreads / kTile; k++) e += sen[k];
        if (e != 0.0) acc_add(energy + (blockIdx.x & (kEnergySlots - 1)), e);
    }
}

// FP32: at most 128 VGPRs, so that four wavefronts fit per SIMD
__global__ __launch_bounds__(kElecThreads) __attribute__((amdgpu_waves_per_eu(4)))
void k_electrostatics_f32(const int4* __restrict__ act, const int* __restrict__ nact, const int* __restrict__ rows,
                          const int* __restrict__ owner,
                          const unsigned long long* __restrict__ tmask, const unsigned long long* __restrict__ tmaskcc,
                          const int* __restrict__ chg, PairAtoms<float> A, const float* __restrict__ mu, PairParams<float> P,
                          const double4* __restrict__ pos64, const double* __restrict__ lab64, PairParams<double> P64,
                          acc_t* __restrict__ force, acc_t* __restrict__ torque, acc_t* __restrict__ energy) {
    electrostatics_tile<float, false>(act, nact, rows, owner, tmask, tmaskcc, chg, A, mu, P, pos64, lab64, P64, force,
                                      torque,
                                      energy, nullptr, 0);
}
__global__ __launch_bounds__(kElecThreads)
void k_electrostatics_f64(const int4* __restrict__ act, const int* __restrict__ nact, const int* __restrict__ rows,
                          const int* __restrict__ owner,
                          const unsigned long long* __restrict__ tmask, const unsigned long long* __restrict__ tmaskcc,
                          const int* __restrict__ chg, PairAtoms<double> A, const double* __restrict__ mu, PairParams<double> P,
                          const double4* __restrict__ pos64, const double* __restrict__ lab64, PairParams<double> P64,
                          acc_t* __restrict__ force, acc_t* __restrict__ torque, acc_t* __restrict__ energy) {
    electrostatics_tile<double, false>(act, nact, rows, owner, tmask, tmaskcc, chg, A, mu, P, pos64, lab64, P64, force,
                                       torque,
                                       energy, nullptr, 0);
}
// The final pass of the energy gradients: the same, with the potential vectors.
template <class R>
__global__ __launch_bounds__(kElecThreads)
void k_electrostatics_pot(const int4* __restrict__ act, const int* __restrict__ nact, const int* __restrict__ rows,
                          const int* __restrict__ owner,
                          const unsigned long long* __restrict__ tmask, const unsigned long long* __restrict__ tmaskcc,
                          const int* __restrict__ chg, PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                          const double4* __restrict__ pos64, const double* __restrict__ lab64, PairParams<double> P64,
                          acc_t* __restrict__ force, acc_t* __restrict__ torque, acc_t* __restrict__ energy,
                          acc_t* __restrict__ pot, int all_general) {
    electrostatics_tile<R, true>(act, nact, rows, owner, tmask, tmaskcc, chg, A, mu, P, pos64, lab64, P64, force, torque,
                                 energy,
                                 pot, all_general);
}
template <class R> struct ElecKernel;
template <> struct ElecKernel<float> {
    static constexpr auto fn = k_electrostatics_f32;
};
template <> struct ElecKernel<double> {
    static constexpr auto fn = k_electrostatics_f64;
};

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
        const int e = A.partners[(size_t)i * A.pstride + k];
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

// pol_only: only pairs of two polarisable sites (the solver's operator).
// GRAD: also the field gradient (xx yy zz xy xz yz) into grad[6 i ..]
// (Extrapolated polarization).
template <class R, bool GRAD>
__global__ __launch_bounds__(256) void k_cov_induced_field(PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                                                           int pol_only, acc_t* __restrict__ field,
                                                           acc_t* __restrict__ grad) {
    if (skipped(P.skip)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int np = A.npart[i];
    if (np == 0) return;
    const R4<R> pi = A.pos[i];
    const V3<R> ui = mk(mu[3 * (size_t)i], mu[3 * (size_t)i + 1], mu[3 * (size_t)i + 2]);
    const R2<R> tdi = A.td[i];
    if (pol_only && tdi.y == R(0)) return;
    R ex = 0, ey = 0, ez = 0;
    R gi[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < np; k++) {
        const int e = A.partners[(size_t)i * A.pstride + k];
        const int j = e >> 2;
        if (j <= i) continue;
        if (pol_only && A.td[j].y == R(0)) continue;
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        const R r2 = dot(dr, dr);
        if (r2 > P.rc2) continue;
        V3<R> fi, fj;
        const R2<R> tdj = A.td[j];
        const R a = partner_scale<R>(e, P.scale14) == R(0) ? tdi.x + tdj.x : P.default_thole;
        const V3<R> uj = mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]);
        if constexpr (GRAD) {
            const InducedRadialGrad<R> f = induced_grad_radial<R>(r2, tdi.y, tdj.y, a, P);
            const InducedRadial<R> f12{f.f1, f.f2};
            fi = induced_field_from<R>(uj, dr, f12);
            fj = induced_field_from<R>(ui, dr, f12);
            // at i from mu_j (separation dr), at j from mu_i (separation -dr)
            induced_grad_add<R>(uj, dr, f.f2, f.f3, gi);
            R gj[6] = {0, 0, 0, 0, 0, 0};
            induced_grad_add<R>(ui, dr * R(-1), f.f2, f.f3, gj);
#pragma unroll
            for (int c = 0; c < 6; c++) acc_add(&grad[6 * (size_t)j + c], gj[c]);
        } else {
            induced_field_pair<R>(dr, r2, ui, uj, tdi.y, tdj.y, a, P, fi, fj);
        }
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
    if constexpr (GRAD) {
#pragma unroll
        for (int c = 0; c < 6; c++) acc_add(&grad[6 * (size_t)i + c], gi[c]);
    }
}

// kPot: also the pairs' potential vectors (dE/dm, 20 per atom) into pot.
template <class R, bool kPot>
__global__ __launch_bounds__(128) void k_cov_electrostatics(PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                                                            acc_t* __restrict__ force, acc_t* __restrict__ torque,
                                                            acc_t* __restrict__ energy, acc_t* __restrict__ pot) {
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
        R pi_[kPot ? kMP : 1] = {};
        bool any = false;
        for (int k = 0; k < np; k++) {
            const int e = A.partners[(size_t)i * A.pstride + k];
            const int j = e >> 2;
            if (j <= i) continue;
            const R4<R> pj = A.pos[j];
            const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
            const R r2 = dot(dr, dr);
            if (r2 > P.rc2) continue;
            const R sc = partner_scale<R>(e, P.scale14);
            R pj_[kPot ? kMP : 1] = {};
            const ElecOut<R> o =
                elec_pair<R, kPot>(dr, r2, mi, A.lab + (size_t)j * kMP, ui,
                                   mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]), tdi, A.td[j],
                                   ani, A.aniso[j] != 0, sc, sc, P, true, pi_, pj_);
            if constexpr (kPot) {
#pragma unroll
                for (int c = 0; c < kMP; c++) acc_add(&pot[kMP * (size_t)j + c], pj_[c]);
            }
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
            if constexpr (kPot) {
#pragma unroll
                for (int c = 0; c < kMP; c++) acc_add(&pot[kMP * (size_t)i + c], pi_[c]);
            }
        }
    }
    // every lane reaches the reduction
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) en += __shfl_xor(en, off);
    if ((threadIdx.x & 63) == 0 && en != 0.0) acc_add(energy + (blockIdx.x & (kEnergySlots - 1)), en);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class R>
void launch_block_bounds(hipStream_t s, int n, const R4<R>* pos, float4* bmin, float4* bmax) {
    int nb = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_block_bounds<R>, dim3(nb), dim3(kTile), 0, s, n, pos, bmin, bmax);
}


template <class R>
void launch_electrostatics_pairs(hipStream_t s, int maxtiles, const int4* act, const int* nact, const int* rows,
                                 const int* owner, const unsigned long long* tmask,
                                 const unsigned long long* tmaskcc, const int* chg, PairAtoms<R> a, const R* mu,
                                 PairParams<R> p, const double4* pos64, const double* lab64, PairParams<double> p64,
                                 acc_t* force, acc_t* torque, acc_t* energy) {
    if (maxtiles <= 0) return;
    hipLaunchKernelGGL(ElecKernel<R>::fn, dim3(maxtiles), dim3(kElecThreads), 0, s, act, nact, rows, owner, tmask, tmaskcc,
                       chg, a, mu, p, pos64, lab64, p64, force, torque, energy);
}

template <class R> void launch_cov_fixed_field(hipStream_t s, int n, PairAtoms<R> a, PairParams<R> p, acc_t* field) {
    hipLaunchKernelGGL(k_cov_fixed_field<R>, dim3((n + 255) / 256), dim3(256), 0, s, a, p, field);
}
template <class R>
void launch_cov_induced_field(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, int pol_only,
                              acc_t* field, acc_t* grad) {
    if (grad)
        hipLaunchKernelGGL((k_cov_induced_field<R, true>), dim3((n + 255) / 256), dim3(256), 0, s, a, mu, p, pol_only,
                           field, grad);
    else
        hipLaunchKernelGGL((k_cov_induced_field<R, false>), dim3((n + 255) / 256), dim3(256), 0, s, a, mu, p,
                           pol_only, field, nullptr);
}
template <class R>
void launch_cov_electrostatics(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, acc_t* force,
                               acc_t* torque, acc_t* energy, acc_t* pot) {
    if (pot)
        hipLaunchKernelGGL((k_cov_electrostatics<R, true>), dim3((n + 127) / 128), dim3(128), 0, s, a, mu, p, force,
                           torque, energy, pot);
    else
        hipLaunchKernelGGL((k_cov_electrostatics<R, false>), dim3((n + 127) / 128), dim3(128), 0, s, a, mu, p, force,
                           torque, energy, nullptr);
}

// ---------------------------------------------------------------------------
// Thole damping of the covalent pairs (energy gradients): h = u dE/du of
// each pair of polarisable partners, once per pair (thread = atom i,
// partners j > i).  hd takes h at both atoms (the damping factors); a pair
// at polarization scale 0 has the width thole_i + thole_j, so ht takes h/a
// at both atoms; any other pair has the default width, and hdef takes 2h
// (hdef sums h over both atoms of every pair, like the neighbour-list kernel).
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void k_cov_thole(PairAtoms<R> A, const R* __restrict__ mu, PairParams<R> P,
                                                   acc_t* __restrict__ hd, acc_t* __restrict__ hdef,
                                                   acc_t* __restrict__ ht) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int np = A.npart[i];
    const R2<R> tdi = A.td[i];
    if (np == 0 || tdi.y == R(0)) return;
    const R4<R> pi = A.pos[i];
    const R* mi = A.lab + (size_t)i * kMP;
    const V3<R> ui = mk(mu[3 * (size_t)i], mu[3 * (size_t)i + 1], mu[3 * (size_t)i + 2]);
    R hi = 0, hdi = 0, hti = 0;
    for (int k = 0; k < np; k++) {
        const int e = A.partners[(size_t)i * A.pstride + k];
        const int j = e >> 2;
        if (j <= i) continue;
        const R2<R> tdj = A.td[j];
        if (tdj.y == R(0)) continue;
        const R4<R> pj = A.pos[j];
        const V3<R> dr = min_image(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), P.box);
        const R r2 = dot(dr, dr);
        if (r2 > P.rc2) continue;
        const R ps = partner_scale<R>(e, P.scale14);
        const R a = ps == R(0) ? tdi.x + tdj.x : P.default_thole;
        const V3<R> uj = mk(mu[3 * (size_t)j], mu[3 * (size_t)j + 1], mu[3 * (size_t)j + 2]);
        const R h = thole_u_derivative<R>(dr, r2, mi, A.lab + (size_t)j * kMP, ui, uj, tdi.y, tdj.y, a, ps,
                                          P.mutual != 0);
        if (h == R(0)) continue;
        hi += h;
        acc_add(&hd[j], h);
        if (ps == R(0)) {
            const R t = h / a;
            hti += t;
            acc_add(&ht[j], t);
        } else {
            hdi += R(2) * h;
        }
    }
    if (hi != R(0)) acc_add(&hd[i], hi);
    if (hti != R(0)) acc_add(&ht[i], hti);
    if (hdi != R(0)) acc_add(&hdef[i], hdi);
}
template <class R>
void launch_cov_thole(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, acc_t* hd, acc_t* hdef,
                      acc_t* ht) {
    hipLaunchKernelGGL(k_cov_thole<R>, dim3((n + 255) / 256), dim3(256), 0, s, a, mu, p, hd, hdef, ht);
}
template <class R>
void launch_electrostatics_pot(hipStream_t s, int maxtiles, const int4* act, const int* nact, const int* rows,
                               const int* owner, const unsigned long long* tmask, const unsigned long long* tmaskcc,
                               const int* chg, PairAtoms<R> a, const R* mu, PairParams<R> p, const double4* pos64,
                               const double* lab64, PairParams<double> p64, acc_t* force, acc_t* torque,
                               acc_t* energy, acc_t* pot, int all_general) {
    if (maxtiles <= 0) return;
    hipLaunchKernelGGL(k_electrostatics_pot<R>, dim3(maxtiles), dim3(kElecThreads), 0, s, act, nact, rows, owner, tmask,
                       tmaskcc, chg, a, mu, p, pos64, lab64, p64, force, torque, energy, pot, all_general);
}

#define MPID_INSTANTIATE(R)                                                                                        \
    template void launch_block_bounds<R>(hipStream_t, int, const R4<R>*, float4*, float4*);                         \
    template void launch_electrostatics_pairs<R>(hipStream_t, int, const int4*, const int*, const int*, const int*,\
                                                 const unsigned long long*, const unsigned long long*, const int*,   \
                                                 PairAtoms<R>, const R*, PairParams<R>, const double4*,             \
                                                 const double*, PairParams<double>, acc_t*, acc_t*, acc_t*);        \
    template void launch_cov_fixed_field<R>(hipStream_t, int, PairAtoms<R>, PairParams<R>, acc_t*);                \
    template void launch_cov_induced_field<R>(hipStream_t, int, PairAtoms<R>, const R*, PairParams<R>, int,        \
                                              acc_t*, acc_t*);                                                    \
    template void launch_cov_electrostatics<R>(hipStream_t, int, PairAtoms<R>, const R*, PairParams<R>, acc_t*,    \
                                               acc_t*, acc_t*, acc_t*);                                             \
    template void launch_cov_thole<R>(hipStream_t, int, PairAtoms<R>, const R*, PairParams<R>, acc_t*, acc_t*,      \
                                      acc_t*);                                                                      \
    template void launch_electrostatics_pot<R>(hipStream_t, int, const int4*, const int*, const int*, const int*,   \
                                               const unsigned long long*, const unsigned long long*, const int*,     \
                                               PairAtoms<R>, const R*, PairParams<R>, const double4*, const double*, \
                                               PairParams<double>, acc_t*, acc_t*, acc_t*, acc_t*, int);
MPID_INSTANTIATE(float)
MPID_INSTANTIATE(double)
#undef MPID_INSTANTIATE

}  // namespace mpid
