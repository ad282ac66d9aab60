// Host-side launch interface of the MPIDForce gfx950 kernels.  Every launcher
// enqueues on the given stream and returns immediately.
#pragma once

#include "common.h"

namespace mpid {

template <class R> struct R4T;
template <> struct R4T<float> { typedef float4 type; };
template <> struct R4T<double> { typedef double4 type; };
template <class R> using R4 = typename R4T<R>::type;

template <class R> struct R2T;
template <> struct R2T<float> { typedef float2 type; };
template <> struct R2T<double> { typedef double2 type; };
template <class R> using R2 = typename R2T<R>::type;

// Grid coordinates u_k = n_k s_k of a position (s its fractional coordinates
// along the box vectors): A_jk = du_k/dr_j = n_k (box^-1)_jk, lower
// triangular (A01 = A02 = A12 = 0); a rectangular box has the diagonal only
// (tri = 0).  Cartesian and fractional multipoles and potential derivatives
// convert through it (transformMultipolesToFractionalCoordinates and
// transformPotentialToCartesianCoordinates, MPIDReferenceForce.cpp:3077-3267).
template <class R> struct FracT {
    R a00, a10, a11, a20, a21, a22;
    int tri;
};
// v_k = sum_j d_j A_jk: the fractional components of a dipole (or of any
// contravariant vector)
template <class R> MPID_HD V3<R> frac_vector(V3<R> d, const FracT<R>& A) {
    return V3<R>{d.x * A.a00 + d.y * A.a10 + d.z * A.a20, d.y * A.a11 + d.z * A.a21, d.z * A.a22};
}
// c_j = sum_k A_jk g_k: Cartesian derivatives from grid-coordinate ones
template <class R> MPID_HD V3<R> cart_vector(V3<R> g, const FracT<R>& A) {
    return V3<R>{A.a00 * g.x, A.a10 * g.x + A.a11 * g.y, A.a20 * g.x + A.a21 * g.y + A.a22 * g.z};
}
// Fractional moments of a Cartesian record (q, dipole, quadrupole XX XY XZ YY
// YZ ZZ, octopole XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ): each tensor
// contracted with A on every index, times the number of index orderings its
// stored component stands for (1 2 2 1 2 1; 1 3 3 3 6 3 1 3 3 1), so that
// sum_k mf_k phi_k over the grid-coordinate derivatives is the energy.
template <class R> MPID_HD void frac_moments(const R* m, const FracT<R>& A, R* mf) {
    mf[0] = m[0];
    if (!A.tri) {
        const R gx = A.a00, gy = A.a11, gz = A.a22;
        mf[1] = gx * m[1]; mf[2] = gy * m[2]; mf[3] = gz * m[3];
        mf[4] = gx * gx * m[4]; mf[5] = R(2) * gx * gy * m[5]; mf[6] = R(2) * gx * gz * m[6];
        mf[7] = gy * gy * m[7]; mf[8] = R(2) * gy * gz * m[8]; mf[9] = gz * gz * m[9];
        mf[10] = gx * gx * gx * m[10]; mf[11] = R(3) * gx * gx * gy * m[11];
        mf[12] = R(3) * gx * gx * gz * m[12]; mf[13] = R(3) * gx * gy * gy * m[13];
        mf[14] = R(6) * gx * gy * gz * m[14]; mf[15] = R(3) * gx * gz * gz * m[15];
        mf[16] = gy * gy * gy * m[16]; mf[17] = R(3) * gy * gy * gz * m[17];
        mf[18] = R(3) * gy * gz * gz * m[18]; mf[19] = gz * gz * gz * m[19];
        return;
    }
    const R a[3][3] = {{A.a00, R(0), R(0)}, {A.a10, A.a11, R(0)}, {A.a20, A.a21, A.a22}};
    const V3<R> d = frac_vector(mk(m[1], m[2], m[3]), A);
    mf[1] = d.x; mf[2] = d.y; mf[3] = d.z;
    // full tensors from the stored components
    const int q2[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    const int q3[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}}, {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
                             {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
    R t2[3][3];  // Q A (second index)
    for (int i = 0; i < 3; i++)
        for (int l = 0; l < 3; l++) {
            R v = R(0);
            for (int j = 0; j < 3; j++) v += m[4 + q2[i][j]] * a[j][l];
            t2[i][l] = v;
        }
    const int kk[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    const R mult2[6] = {R(1), R(2), R(2), R(1), R(2), R(1)};
    for (int c = 0; c < 6; c++) {
        const int k = kk[c][0], l = kk[c][1];
        R v = R(0);
        for (int i = 0; i < 3; i++) v += a[i][k] * t2[i][l];
        mf[4 + c] = mult2[c] * v;
    }
    R t3[3][3][3];  // O contracted with A on its last index
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int n = 0; n < 3; n++) {
                R v = R(0);
                for (int p = 0; p < 3; p++) v += m[10 + q3[i][j][p]] * a[p][n];
                t3[i][j][n] = v;
            }
    const int kkk[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                            {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
    const R mult3[10] = {R(1), R(3), R(3), R(3), R(6), R(3), R(1), R(3), R(3), R(1)};
    for (int c = 0; c < 10; c++) {
        const int k = kkk[c][0], l = kkk[c][1], n = kkk[c][2];
        R v = R(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) v += a[i][k] * a[j][l] * t3[i][j][n];
        mf[10 + c] = mult3[c] * v;
    }
}

// Parameters of the real-space pair kernels.
template <class R> struct PairParams {
    Box<R> box;
    int n;
    R rc2;
    R alpha;
    R aln[5];          // (2 alpha^2)^k / (alpha sqrt(pi)), k = 1..4 (index 0 unused)
    R default_thole;
    R scale14;
    int mutual;        // 1 when the induced-induced terms enter the forces
    int fix_torque;    // 1: energy-conserving induced-induced torques (MPID_FLAG_CONSERVATIVE_TORQUES)
    const double* skip;  // the solver's stop flag in the induced-field kernels of a solver iteration, else null
};

// Precision-specific per-atom inputs of the pair kernels (sorted order).
template <class R> struct PairAtoms {
    const R4<R>* pos;      // x, y, z, damping factor
    const R* lab;          // [n*20] lab-frame multipoles
    const R2<R>* td;       // thole, damping factor
    const int* aniso;      // [n]
    const int* partners;   // [n*pstride] (sorted index << 2 | class)
    const int* npart;      // [n]
    int pstride;           // partner slots per atom: the largest partner count of the force
};

constexpr int kNlSplit = 4;  // threads per atom in the list-driven kernels
// Per-atom neighbour lists (kernels_nlist.hip).  Entry k of list row l sits at
// ((k / kNlSplit) * rows + l) * kNlSplit + k % kNlSplit; a kernel given an
// atoms index reads row t for atoms[t] (rows = the index's length), one
// without reads row i for atom i (rows = n).
struct NbrList {
    const int* nbr;    // sorted indices of the neighbours
    const int* count;  // [n], by atom index
};

// ---- atoms / frames / solver vectors (kernels_atoms.hip) ---------------------
void launch_morton_keys(hipStream_t s, int n, const double* pos, Box<double> box, unsigned* keys, int* vals);
void launch_invert_perm(hipStream_t s, int n, const int* perm, int* inv);
void launch_gather_sorted(hipStream_t s, int n, const int* perm, const int* inv, const double* pos, Box<double> box,
                          const double* mol, const int4* axes, const double* thole, const double* pol,
                          const int* partners, const int* npart, int pstride, double4* spos, double* smol,
                          int4* saxes, double* sthole, double* spol, int* spartners, int* snpart);
// chg[s] = 1: the site carries a charge only (no higher moment, no polarisability)
void launch_rotate(hipStream_t s, int n, Box<double> box, const double4* pos, const int4* axes, const double* mol,
                   const double* pol, double* lab, double* labpol, double* sqrtpol, double* damp, int* aniso,
                   int* chg);
template <class R>
void launch_convert(hipStream_t s, int n, const double4* pos, const double* lab, const double* thole,
                    const double* damp, R4<R>* rpos, R* rlab, R2<R>* rtd);
// exact_threefold: ThreeFold frames by the exact chain rule (conservative
// torques) instead of the reference's approximate projection
void launch_torque_to_force(hipStream_t s, int n, Box<double> box, const double4* pos, const int4* axes,
                            const acc_t* torque, acc_t* force, int exact_threefold);
void launch_output_forces(hipStream_t s, int n, const int* perm, const acc_t* force, double* out, int accumulate);
void launch_unpermute3(hipStream_t s, int n, const int* perm, const double* in, double* out);
void launch_zero_acc(hipStream_t s, acc_t* p, size_t count, const double* skip = nullptr);
size_t sort_temp_bytes(int n);
int launch_sort_pairs(hipStream_t s, void* temp, size_t bytes, unsigned* kin, unsigned* kout, int* vin, int* vout, int n);
// out = acc + erecip + selfc * mu (field of the induced dipoles, double)
template <class R>
void launch_field_sum(hipStream_t s, int n, const acc_t* acc, const R* erec, const R* mu, double selfc, double* out);
template <class R> void launch_fill(hipStream_t s, R* p, size_t count, R v);

// fixed field: E0 = recip (from phi) + self + real-space accumulator
template <class R>
void launch_fixed_field_finish(hipStream_t s, int n, const R* phi, const double* lab, const acc_t* acc,
                               FracT<double> A, double selfc, double* E0);
// mu0 = alpha_lab E0 ; y0 = S E0 ; b = S E0
void launch_solver_init(hipStream_t s, int n, const double* sqrtpol, const double* E0, double* b, double* y);
// mu = S y (double and precision copies)
template <class R> void launch_apply_S(hipStream_t s, int n, const double* sqrtpol, const double* y, R* mu, double* mud);
// Ap = p - S (acc + Erecip + selfc mu)
template <class R>
void launch_matvec_finish(hipStream_t s, int n, const double* sqrtpol, const acc_t* acc, const R* erecip,
                          const R* mu, double selfc, const double* p, double* Ap);
// deterministic FP64 reductions
void launch_dot(hipStream_t s, int m, const double* a, const double* b, double* partial, double* result);
// Getters after an evaluation (sorted-order state).  Potential at npts points
// of the charges, dipoles plus induced dipoles and quadrupoles (FP64).
void launch_potential_points(hipStream_t s, int n, int npts, const double4* pos, const double* lab, const double* mu,
                             Box<double> box, const double* points, double* out);
// The allocation-free form of the potential getter stages its points and
// values in a file-scope device array of kNoAllocScratchDoubles doubles
// (kernels_atoms.hip): at most kNoAllocMaxPoints points per call.  Its
// address on the current device (looked up once per device).
constexpr int kNoAllocMaxPoints = 65536;
constexpr int kNoAllocScratchDoubles = 4 * kNoAllocMaxPoints;
hipError_t noalloc_scratch_address(double** out);
// System multipole sums (one block): out[0] total mass, [1..3] mass-weighted
// position, [4] charge, [5..7] dipole, [8..13] traced second moment about the
// centre of mass (xx xy xz yy yz zz), [14..19] atomic quadrupole sum.
void launch_multipole_moments(hipStream_t s, int n, const int* perm, const double* pos, const double* masses,
                              const double* lab, const double* mu, double* out);
// Energies accumulate in fixed point over kEnergySlots words per term (a
// workgroup adds to slot blockIdx % kEnergySlots; the host sums the slots):
// atomics from every workgroup on one word serialise across the chip.
constexpr int kEnergySlots = 256;
// Extrapolated (OPT) polarization: the partial sums c'_o = sum_{j>=o} c_j of
// the coefficients (at most kMaxExtrapolation), passed by value.
constexpr int kMaxExtrapolation = 16;
struct ExtrapCoefficients {
    int n;
    double c[kMaxExtrapolation];
};
// One order: E_(m), G_(m) (field with the self term, gradient without it, of
// mu_(m) at every atom, from the real-space accumulators, the reciprocal
// potential derivatives phidp and the self term) and mu_(m+1) = alpha_lab E.
template <class R>
void launch_extrap_order(hipStream_t s, int n, const acc_t* field, const acc_t* gacc, const R* phidp, FracT<double> A,
                         double selfc, const double* labpol, const double* mu, double* E, double* G, double* mu_next,
                         R* mu_next_r);
// mu = sum_o c'_o mu_(o) (mus: [order][3n])
template <class R>
void launch_extrap_combine(hipStream_t s, int n, const ExtrapCoefficients& cp, const double* mus, double* mud,
                           R* mu_r);
// the dipole-response forces and torques into the fixed-point accumulators
void launch_extrap_response(hipStream_t s, int n, const ExtrapCoefficients& cp, const double* mus, const double* Es,
                            const double* Gs, const int* aniso, acc_t* force, acc_t* torque);
// The solver's device-side state in scal: [0] rr, [1] pAp, [2] rr_new,
// [3] |S r|^2, [7] beta, [kCgCount] iterations done, [kCgStop] 1 once the
// stopping test has passed.  Kernels of an iteration given the stop flag
// return at once when it is set, so that the host may issue an iteration
// before it has read the previous one's check.
constexpr int kCgCount = 9;
constexpr int kCgStop = 10;
constexpr int kScalars = 16;
// Check 0 of a solve: the stopping test on the initial |S r|^2 (scal[3]),
// published to scal and to hcheck (pinned, mapped host memory, two slots of
// (|S r|^2, stop); check c goes to slot c & 1).
void launch_cg_begin(hipStream_t s, int n, double* scal, double* hcheck, double stop_eps);
// One CG iteration once the operator's field for mu = S p is complete (acc,
// erec): Ap, a = rr / pAp, x += a p, r -= a Ap, the norms rr and |S r|^2,
// beta, the iteration's check, p = r + beta p and the next operator input
// mu = S p.
template <class R>
void launch_cg_iteration(hipStream_t s, int n, const double* sqrtpol, const acc_t* acc, const R* erec, R* mu,
                         double selfc, double* p, double* Ap, double* x, double* r, double* partial, double* scal,
                         double* hcheck, double stop_eps);
void launch_residual_norm(hipStream_t s, int n, const double* sqrtpol, const double* r, double* partial,
                          double* result);
void launch_axpy(hipStream_t s, int m, double a, const double* x, double* y);
void launch_sub(hipStream_t s, int m, const double* a, const double* b, double* out);
// reciprocal-space energy, forces and torques, self terms (per atom)
template <class R>
void launch_recip_atoms(hipStream_t s, int n, int mutual, const double* phi, const R* phidp, const double* lab,
                        const double* mu, const int* aniso, FracT<double> A, double alpha, acc_t* force,
                        acc_t* torque, acc_t* energy);

// ---- PME (kernels_pme.hip) ----------------------------------------------------
template <class R> struct PmeGeom {
    int n[3];
    R gscale[3];      // n_k / L_k (the diagonal of A)
    FracT<R> A;
    Box<R> box;
    const double* skip;  // the solver's stop flag in the spreading and gathering of a solver iteration, else null
};
// Morton-aligned spreading tiles: 2^(3m) tiles of T_k grid points, windows of W_k points
struct TileGeom {
    int m;               // Morton level
    int tiles_per_axis;  // 2^m
    int ntiles;          // 2^(3m)
    int T[3];
    int W[3];
};
// Morton-cell ranges of the sorted atoms, or (idx != nullptr) of the subsequence idx[0 .. *count)
void launch_tile_ranges(hipStream_t s, int n, const unsigned* keys_sorted, const int* idx, const int* count,
                        int shift, int* tstart, int* tend);
size_t spread_tiled_lds_bytes(const TileGeom& tg, int double_precision);
template <class R>
void launch_spread_fixed_tiled(hipStream_t s, const R4<R>* pos, const R* lab, PmeGeom<R> g, TileGeom tg,
                               const int* tstart, const int* tend, R* scratch, R* grid);
// atoms != nullptr: the ranges index atoms[] (the polarisable subsequence)
template <class R>
void launch_spread_induced_tiled(hipStream_t s, const R4<R>* pos, const R* mu, PmeGeom<R> g, TileGeom tg,
                                 const int* atoms, const int* tstart, const int* tend, R* scratch, R* grid);
template <class R> void launch_fold(hipStream_t s, PmeGeom<R> g, TileGeom tg, const R* scratch, R* grid);
// Owner-computes spreading (tiles of at least 8 points per axis): one
// workgroup per grid core writes the whole grid, no scratch and no fold.
bool spread_by_owner(const TileGeom& tg);
template <class R>
void launch_spread_fixed_owner(hipStream_t s, const R4<R>* pos, const R* lab, PmeGeom<R> g, TileGeom tg,
                               const int* tstart, const int* tend, R* grid);
template <class R>
void launch_spread_induced_owner(hipStream_t s, const R4<R>* pos, const R* mu, PmeGeom<R> g, TileGeom tg,
                                 const int* atoms, const int* tstart, const int* tend, R* grid);
template <class R>
void launch_spread_fixed(hipStream_t s, int n, const R4<R>* pos, const R* lab, PmeGeom<R> g, R* grid);
template <class R>
void launch_spread_induced(hipStream_t s, int n, const R4<R>* pos, const R* mu, PmeGeom<R> g, R* grid);
template <class R>
void launch_gather_phi(hipStream_t s, int n, const R4<R>* pos, PmeGeom<R> g, const R* grid, R* phi);
template <class R>
void launch_gather_field(hipStream_t s, int natoms, const int* atoms, const R4<R>* pos, PmeGeom<R> g, const R* grid,
                         R* field);
// ---- FFT and convolution (kernels_fft.hip) ------------------------------------
// x / d = (x * m) >> 40 with m = ceil(2^40 / d), exact for x < 2^20.
struct FastDiv {
    unsigned d;
    unsigned long long m;
};
FastDiv make_fastdiv(unsigned d);
__device__ __forceinline__ unsigned fdiv(unsigned x, FastDiv f) {
    return (unsigned)(((unsigned long long)x * f.m) >> 40);
}
constexpr int kFftMaxStages = 8;
struct FftStage {
    int rad;      // 4, 2, 3 or 5
    FastDiv nb;   // n / rad
    FastDiv ns;   // product of the earlier radices
    int twstep;   // n / (ns * rad)
};
struct FftAxis {
    int n, nst;
    FastDiv nd;
    FftStage st[kFftMaxStages];
};
bool make_fft_axis(int n, FftAxis& ax);  // false unless n = 2^a 3^b 5^c with at most 8 stages
template <class R> struct FftPlan {
    FftAxis ax[3];
    const R2<R>* tw[3];  // per-axis twiddles exp(-2 pi i k / n), k < n
};
template <class R> struct ConvParams {
    const R* infl;       // the influence function on the half spectrum, [kz][ky][kx] (launch_influence_table)
    const double* skip;  // the solver's stop flag in the FFT passes of a solver iteration, else null
};
// Inputs of the influence-function table, FP64
struct InflParams {
    int n[3];
    double expf;                          // pi^2 / alpha^2
    double sc;                            // 1 / (pi V)
    double r00, r10, r11, r20, r21, r22;  // box^-1 (lower triangular): m = box^-1 k
};
// out[(kz * n1 + ky) * n0 + kx], kz < n2 / 2 + 1; moduli64: |b(k)|^2, [n0 | n1 | n2]
template <class R> void launch_influence_table(hipStream_t s, InflParams ip, const double* moduli64, R* out);
// Fold for tilings with W = T + 8 <= 2T in every dimension: one workgroup per
// Morton cell sums, for each point of its T^3 core, its own window and those
// of its lower neighbours (launch_fold handles any tiling)
bool fold_by_tiles(const TileGeom& tg);
template <class R> void launch_fold_tiles(hipStream_t s, PmeGeom<R> g, TileGeom tg, const R* scratch, R* grid);
// Q (real grid) -> S (half spectrum) -> convolution -> Q (potential grid)
// (every pass returns at once when cp.skip is set)
template <class R>
void launch_recip_convolution(hipStream_t s, const FftPlan<R>& p, R* Q, R2<R>* S, ConvParams<R> cp);

// ---- real space (kernels_pairs.hip) -------------------------------------------
template <class R>
void launch_block_bounds(hipStream_t s, int n, const R4<R>* pos, float4* bmin, float4* bmax);
// One workgroup per tile act[0 .. *nact) (row entries r of block pairs
// J >= I); the launch covers maxtiles >= *nact.
// tmask: pairs with a multipole or a polarisable site; tmaskcc: pairs of two
// bare charges; chg[i] != 0: atom i is a bare charge (pairs of tmask with one
// such site take the reduced formula).  The charge-charge channel of all
// pairs is evaluated in FP64 from pos64 (x, y, z) and the FP64 lab-frame
// charges lab64[20 i].
template <class R>
void launch_electrostatics_pairs(hipStream_t s, int maxtiles, const int4* act, const int* nact, const int* rows,
                                 const int* owner, const unsigned long long* tmask,
                                 const unsigned long long* tmaskcc, const int* chg, PairAtoms<R> a, const R* mu,
                                 PairParams<R> p, const double4* pos64, const double* lab64, PairParams<double> p64,
                                 acc_t* force, acc_t* torque, acc_t* energy);
// ---- neighbour lists (kernels_nlist.hip) ---------------------------------------
// Block pairs within the cutoff: counts and upper (those with J >= I) when
// rows == nullptr, or the rows at start[I] (at most cap entries are stored;
// the count may exceed it)
void launch_block_rows(hipStream_t s, int nb, const float4* bmin, const float4* bmax, Box<float> box, float cutoff,
                       const int* start, int* counts, int* upper, int* rows, int* owner, int cap);
void launch_exclusive_scan(hipStream_t s, int m, const int* counts, int* start);
// pos.w of the FP64 positions is the damping factor: nonzero iff the site is polarisable
// Masks, then lists: all = true gives the full list of every atom stored by
// atom index; all = false the full and polarisable-neighbour lists of the
// polarisable atoms, stored by position in their sorted index (polstart,
// the exclusive scan of polcnt, is written on the way).  nmask: 64 words per
// row entry (nrows entries); polmask: one word per block; scalars[0]:
// largest stored list length, scalars[1]: nonzero if a block pair lacked its
// mirror entry.  act[ustart[I] + k] = {r, I, J} for the k-th row entry r of
// row I with J >= I, r = -1 if the block pair holds no pair (ustart: the
// exclusive scan of launch_block_rows' upper, whose last element is the
// number of such entries).  nrows: row entries launched (the count, or the capacity
// when the count is not known on the host: entries past rstart[nb] return);
// rcap: the capacity of rows / nmask (entries beyond it are ignored).
void launch_nl_build(hipStream_t s, int n, int nrows, int rcap, bool all, const int* rstart, const int* rows,
                     const int* owner, const float4* bmin, const float4* bmax, const double4* pos,
                     const int* partners, const int* npart, int pstride, Box<double> box, double rc2, int cap,
                     unsigned long long* nmask, unsigned long long* polmask, int* polcnt, int* polstart, int* nbr,
                     int* count, int* scalars, unsigned long long* tmask, int* pnbr, int* pcount, const int* chg,
                     unsigned long long* tmaskcc, int4* act, const int* ustart);
void launch_pol_fill(hipStream_t s, int n, const double4* pos, const int* start, int* idx);
// fields at the atoms atoms[0..natoms) (all atoms when atoms == nullptr, natoms = n)
template <class R>
void launch_nl_fixed_field(hipStream_t s, int natoms, const int* atoms, NbrList l, PairAtoms<R> a, PairParams<R> p,
                           acc_t* field);
// grad != nullptr: also the field gradient (xx yy zz xy xz yz per atom)
template <class R>
void launch_nl_induced_field(hipStream_t s, int natoms, const int* atoms, NbrList l, PairAtoms<R> a, const R* mu,
                             PairParams<R> p, acc_t* field, acc_t* grad = nullptr);
// the scaled covalent pairs, which the list and tile kernels leave out
template <class R> void launch_cov_fixed_field(hipStream_t s, int n, PairAtoms<R> a, PairParams<R> p, acc_t* field);
template <class R>
void launch_cov_induced_field(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, int pol_only,
                              acc_t* field, acc_t* grad = nullptr);
// pot != nullptr: also each pair's potential vectors dE/dm (20 per atom)
template <class R>
void launch_cov_electrostatics(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, acc_t* force,
                               acc_t* torque, acc_t* energy, acc_t* pot = nullptr);

// ---- energy gradients ------------------------------------------------------------
// The final pass of the gradients: launch_electrostatics_pairs plus the
// potential vectors dE/dm of every pair (20 stored lab components per atom,
// induced dipoles fixed) into pot.  all_general: every pair takes the general
// formula, so that bare-charge sites get all 20 components (else the charge
// component only).
template <class R>
void launch_electrostatics_pot(hipStream_t s, int maxtiles, const int4* act, const int* nact, const int* rows,
                               const int* owner, const unsigned long long* tmask, const unsigned long long* tmaskcc,
                               const int* chg, PairAtoms<R> a, const R* mu, PairParams<R> p, const double4* pos64,
                               const double* lab64, PairParams<double> p64, acc_t* force, acc_t* torque,
                               acc_t* energy, acc_t* pot, int all_general);
// Thole damping derivatives h = u dE/du (thole_u_derivative) of the pairs of
// polarisable sites.  hd[i]: sum of h over the pairs of i (dE/ddamp_i =
// -hd_i / damp_i); hdef[i]: the same over the pairs at the default Thole
// width, so that dE/d(default width) = sum_i hdef_i / (2 width); ht[i]: sum
// of h / a over the pairs whose width a is thole_i + thole_j (dE/dthole_i).
// List kernel: the polarisable atoms atoms[0..natoms) and their polarisable
// neighbours (each pair from both sides); covalent kernel: the scaled
// partner pairs, each once.
template <class R>
void launch_nl_thole(hipStream_t s, int natoms, const int* atoms, NbrList l, PairAtoms<R> a, const R* mu,
                     PairParams<R> p, acc_t* hd, acc_t* hdef);
template <class R>
void launch_cov_thole(hipStream_t s, int n, PairAtoms<R> a, const R* mu, PairParams<R> p, acc_t* hd, acc_t* hdef,
                      acc_t* ht);
// Reciprocal and self parts of the potential vectors (kernels_atoms.hip);
// pme = 0 leaves out the reciprocal part (NoCutoff, where alpha = 0 removes
// the self part too).
template <class R>
void launch_recip_potential(hipStream_t s, int n, int pme, const double* phi, const R* phidp, const double* lab,
                            const double* mu, FracT<double> A, double alpha, acc_t* pot);
// Total field E0 + T mu (mutual != 0) or E0: the real-space part of T mu in
// acc, its reciprocal part from phidp, the self term selfc mu.
template <class R>
void launch_total_field(hipStream_t s, int n, int mutual, const double* E0, const acc_t* acc, const R* phidp,
                        FracT<double> A, const double* mu, double selfc, double* out);
// The per-atom parameter gradients (caller's order): gmol [20 n] (molecular
// record, internal component order), gpol [3 n], gthole [n]; *gdef += the
// default Thole width's.
void launch_param_gradients(hipStream_t s, int n, int all_sites, const int* perm, Box<double> box, const double4* pos,
                            const int4* axes, const double* mol, const double* pol, const double* damp,
                            const int* chg, const acc_t* pot, const double* Etot, const acc_t* hd,
                            const acc_t* hdef, const acc_t* ht, double default_thole, double* gmol, double* gpol,
                            double* gthole, acc_t* gdef);

}  // namespace mpid
