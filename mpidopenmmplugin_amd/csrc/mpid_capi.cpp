// Host orchestration and C ABI of the MPIDForce gfx950 evaluator.
//
// One mpid_context plays the role of the reference's platform kernel object
// (CalcMPIDForceKernel, openmmapi/include/openmm/mpidKernels.h:50-104):
// mpid_create ~ initialize, mpid_execute ~ execute.  The evaluation sequence
// follows MPIDReferencePmeForce::calculateForceAndEnergy
// (platforms/reference/src/SimTKReference/MPIDReferenceForce.cpp:2193-2270):
// lab-frame rotation, permanent field (reciprocal + self + real space),
// induced dipoles, then real-space and reciprocal energies, forces and torques,
// and finally the torque-to-force chain rule.
//
// Device data live in Morton-sorted atom order for the whole evaluation; the
// permutation is rebuilt from the positions on every call.  The mutual
// induced dipoles are found by conjugate gradients on the symmetrised
// operator I - S T S (S = alpha^(1/2) per site), which has the same solution
// as the reference's DIIS iteration and is stopped by the same RMS-Debye
// criterion (MPIDReferenceForce.cpp:1219).
#include "engine.h"
#include "../../include/mpid_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

using namespace mpid;

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct MpidError : std::runtime_error {
    int code;
    MpidError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(call)                                                                                         \
    do {                                                                                                        \
        hipError_t e_ = (call);                                                                                 \
        if (e_ != hipSuccess)                                                                                   \
            throw HipError(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" +        \
                           std::to_string(__LINE__));                                                           \
    } while (0)

// Owning device buffer.
// MPID_DEBUG_POISON set in the environment: every new device allocation of a
// floating-point type is filled with NaN bytes, so that a kernel reading
// memory that nothing wrote shows up as NaN in the results.
inline bool debug_poison() {
    static const bool on = std::getenv("MPID_DEBUG_POISON") != nullptr;
    return on;
}

template <class T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void resize(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
        if (!std::is_integral<T>::value && debug_poison()) HIP_CHECK(hipMemset(p, 0xFF, count * sizeof(T)));
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        resize(count);
        HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// B-spline moduli |b(k)|^2 with the Euler-spline fix and the optimal-zeta
// correction (initializeBSplineModuli, MPIDReferenceForce.cpp:2720-2810).
std::vector<double> bspline_moduli(int n) {
    // order-6 cardinal B-spline at the integers 1..5 from the Cox-de Boor recursion at w = 0
    double c[7][7] = {};
    c[1][0] = 1.0;
    const double w = 0.0;
    for (int p = 2; p <= kOrder; p++)
        for (int m = 0; m < p; m++) {
            double left = m >= 1 ? c[p - 1][m - 1] : 0.0, right = m <= p - 2 ? c[p - 1][m] : 0.0;
            c[p][m] = ((w + (p - 1 - m)) * left + (m + 1 - w) * right) / (p - 1);
        }
    std::vector<double> b(std::max(n, kOrder + 1), 0.0), mod(n);
    for (int m = 1; m <= kOrder; m++) b[m] = c[kOrder][m - 1];
    for (int i = 0; i < n; i++) {
        double s1 = 0, s2 = 0;
        for (int m = 0; m < n; m++) {
            double arg = 2.0 * M_PI * i * m / n;
            s1 += b[m] * std::cos(arg);
            s2 += b[m] * std::sin(arg);
        }
        mod[i] = s1 * s1 + s2 * s2;
    }
    const double eps = 1e-7;
    if (mod[0] < eps) mod[0] = 0.5 * mod[1];
    for (int i = 1; i < n - 1; i++)
        if (mod[i] < eps) mod[i] = 0.5 * (mod[i - 1] + mod[i + 1]);
    if (mod[n - 1] < eps) mod[n - 1] = 0.5 * mod[n - 2];
    for (int i = 1; i <= n; i++) {
        int k = i - 1;
        if (i > n / 2) k -= n;
        double zeta = 1.0;
        if (k != 0) {
            double s1 = 1.0, s2 = 1.0, f = M_PI * k / n;
            for (int j = 1; j <= 50; j++) {
                double a = f / (f + M_PI * j);
                s1 += std::pow(a, kOrder);
                s2 += std::pow(a, 2 * kOrder);
            }
            for (int j = 1; j <= 50; j++) {
                double a = f / (f - M_PI * j);
                s1 += std::pow(a, kOrder);
                s2 += std::pow(a, 2 * kOrder);
            }
            zeta = s2 / s1;
        }
        mod[i - 1] *= zeta * zeta;
    }
    return mod;
}

bool factor_radices(int n, int* out, int& count) {
    count = 0;
    while (n % 4 == 0) { out[count++] = 4; n /= 4; }
    while (n % 2 == 0) { out[count++] = 2; n /= 2; }
    while (n % 3 == 0) { out[count++] = 3; n /= 3; }
    while (n % 5 == 0) { out[count++] = 5; n /= 5; }
    return n == 1 && count <= 8;
}

enum ProfCat {
    P_SORT = MPID_PROF_SORT, P_ROTATE = MPID_PROF_ROTATE, P_TILES = MPID_PROF_TILES, P_SPREAD = MPID_PROF_SPREAD,
    P_FFT = MPID_PROF_FFT, P_CONV = MPID_PROF_CONVOLVE, P_GATHER = MPID_PROF_GATHER,
    P_FIXED = MPID_PROF_FIXED_PAIRS, P_INDUCED = MPID_PROF_INDUCED_PAIRS, P_ELEC = MPID_PROF_ELECTROSTATIC_PAIRS,
    P_RECIP = MPID_PROF_RECIP_ATOMS, P_SOLVER = MPID_PROF_SOLVER, P_OTHER = MPID_PROF_OTHER,
    P_COV = MPID_PROF_COVALENT_PAIRS, P_REFINE = MPID_PROF_REFINEMENT
};

// Precision-specific device state.
template <class R> struct Prec {
    DBuf<R4<R>> pos;
    DBuf<R> lab;
    DBuf<R2<R>> td;
    DBuf<R> mu;       // induced dipoles in this precision (matvec input)
    DBuf<R> erec;     // reciprocal induced field
    DBuf<R> phi, phidp;
    DBuf<R> grid;       // real charge / potential grid [n0][n1][n2]
    DBuf<R2<R>> spec;   // half spectrum [n0][n1][n2/2+1]
    DBuf<double> moduli;  // |b(k)|^2 in FP64, [n0 | n1 | n2]
    DBuf<R> infl;         // the influence function on the half spectrum (launch_influence_table)
    double infl_key[10] = {};  // box vectors and alpha it was computed for
    bool infl_valid = false;
    DBuf<R2<R>> tw[3];
    DBuf<R> scratch;  // spreading windows of the Morton tiles
    FftPlan<R> plan;
    bool ready = false;
};

}  // namespace

// ---- tilings of the PME spreading ---------------------------------------------
// (level, mode) as mpid_options takes them: level -1 without tiles; level m
// >= 1 cuts each axis of the grid into 2^m tiles of T_k = n_k / 2^m >= 6
// points, spread by owned cores (T_k >= 8 and the core's LDS within 64 KiB)
// or by windows of T_k + 8 points folded into the grid (at most 3 windows
// per point and axis, and the window's LDS within 64 KiB).
enum { kSpreadUntiled = 0, kSpreadOwner = 1, kSpreadWindows = 2 };
struct SpreadTiling {
    int level, mode;
};
static bool tile_geometry(const int grid[3], int m, TileGeom& tg) {
    const int a = 1 << m;
    tg = TileGeom{};
    tg.m = m;
    tg.tiles_per_axis = a;
    tg.ntiles = a * a * a;
    for (int k = 0; k < 3; k++) {
        if (grid[k] % a != 0 || grid[k] / a < 6) return false;
        tg.T[k] = grid[k] / a;
        tg.W[k] = std::min(tg.T[k] + 8, grid[k]);
    }
    return true;
}
static bool windows_fit(const TileGeom& tg) {
    for (int k = 0; k < 3; k++)
        if ((tg.W[k] - 1) / tg.T[k] + 1 > 3) return false;  // the fold sums at most 3 windows per axis
    return spread_tiled_lds_bytes(tg, 1) <= 65536;
}
// Every available tiling; first the built-in choice: the largest level at
// which one of the two tiled forms fits, by owned cores where they fit.
static std::vector<SpreadTiling> spread_tilings(const int grid[3]) {
    std::vector<SpreadTiling> out;
    for (int m = 10; m >= 1; m--) {
        TileGeom tg;
        if (!tile_geometry(grid, m, tg)) continue;
        if (spread_by_owner(tg)) out.push_back(SpreadTiling{m, kSpreadOwner});
        if (windows_fit(tg)) out.push_back(SpreadTiling{m, kSpreadWindows});
    }
    out.push_back(SpreadTiling{-1, kSpreadUntiled});
    return out;
}

struct mpid_context {
    int n = 0;
    mpid_options opt{};
    hipStream_t stream = nullptr;
    // The reciprocal-space chains (spread, FFT, gather) run on rstream beside
    // the real-space kernels on stream: their launches hold few workgroups and
    // leave most CUs free.  fork() makes rstream wait for the work issued to
    // stream so far; join() makes stream wait for the work issued to rstream.
    hipStream_t rstream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // The mixed path's FP64 check of the solution (the true residual) runs on
    // vstream beside the final pass, which needs only the solution itself.
    hipStream_t vstream = nullptr;
    hipEvent_t ev_v = nullptr, ev_vdone = nullptr;
    bool final_issued = false;  // the final pass of the current solution is on its way
    bool defer_final = false;   // the solve leaves the final pass to its caller (energy gradients)
    // CG iterations are replayed from a graph captured once per solve (its
    // kernel arguments are those of that solve); hscal is the pinned host
    // copy of the residual norm that each replay leaves behind.
    hipGraphExec_t cg_exec = nullptr;
    double* hscal_dev = nullptr;  // device address of hscal (pinned, mapped)
    double* hscal = nullptr;  // [1] read_scalar, [2] energy-correction dot, [3] (as int) the block-row count,
                              // [kCheck ..] the solver's two check slots (|S r|^2, stop)
    static constexpr int kCheck = 8;
    hipEvent_t ev_check[2] = {nullptr, nullptr};  // recorded after the work that publishes check c (c & 1)
    hipEvent_t ev_sync = nullptr;
    std::string error;
    mpid_stats stats{};
    // this device's copy of the allocation-free getter's staging array
    double* noalloc_scratch = nullptr;

    // static parameters (original order)
    DBuf<double> mol, thole, pol;
    DBuf<int4> axes;
    DBuf<int> partners, npart;
    int pstride = 1;  // partner slots per atom: the largest scaled-partner count
    std::vector<double> h_charges;

    // sorted per-evaluation state
    DBuf<double> pos_in;
    DBuf<unsigned> keys, keys_sorted;
    DBuf<int> vals, perm, inv;
    DBuf<unsigned char> sort_tmp;
    DBuf<double4> spos;
    DBuf<double> smol, sthole, spol, lab, labpol, sqrtpol, damp;
    DBuf<int4> saxes;
    DBuf<int> spart, snpart, aniso, chg;
    // Morton-aligned PME spreading tiles
    TileGeom tgeom{};
    bool tiled_spread = false;
    bool owner_spread = false;  // tiles of at least 8 points: owner-computes spreading, no scratch or fold
    DBuf<int> tstart, tend;     // over all sorted atoms
    DBuf<int> tpstart, tpend;   // over the polarisable atoms (indices into pol_idx)
    // 64-atom blocks: bounding boxes, block rows (CSR: for each block I the
    // blocks J within the cutoff, with their owner I), per-atom neighbour lists
    // of capacity nl_cap, and the tile masks of the row entries with J >= I
    DBuf<float4> bmin, bmax;
    DBuf<int> brow_count, brow_start, brows, brow_owner, nbr, ncount, nl_scalars;
    DBuf<int4> tile_act;  // the electrostatics kernels' tiles {r, I, J} (row entries with J >= I)
    DBuf<int> brow_upper, brow_ustart;  // per block row: entries with J >= I, and their exclusive scan
    // polarisable neighbours only (pnbr/pcount), and the sorted indices of the polarisable atoms
    DBuf<int> pnbr, pcount, polcnt, polstart, pol_idx;
    int npol = 0;
    DBuf<unsigned long long> tmask, tmaskcc, nmask, polmask;
    int rows_total = 0;
    bool rows_pending = false;  // the block-row count of this build is still on its way to hnl[2]
    int rows_grow_to = 0;       // a row count that overflowed the capacity (grown before the retry)
    bool all_lists = false;  // lists of the last build: every atom (true) or the polarisable ones
    int nl_cap = 0;
    int nl_max = 0;
    // After one verified build for the current parameters, the list build of
    // an evaluation is not waited for: npol is kept (it depends on the
    // polarisabilities only), and the overflow and mirror checks, copied to
    // hnl, are read at the solver's first synchronisation (check_lists_pending).
    int npol_cached = -1;
    bool lists_pending = false;
    int* hnl = nullptr;  // pinned: [0] largest list length, [1] mirror flag
    // accumulators
    DBuf<acc_t> field_acc, force_acc, torque_acc, energy_acc;
    // Extrapolated polarization: coefficient partial sums, the order dipoles
    // mu_(m) [N][3n], the fields E_(m) and gradients G_(m) of mu_(0..N-2)
    // ([N-1][3n], [N-1][6n]) and the real-space gradient accumulator
    ExtrapCoefficients ext_cp{};
    DBuf<double> ptmu, ptE, ptG;
    DBuf<acc_t> grad_acc;
    bool mu_set = false;  // the induced dipoles are in place (Extrapolated); the final pass does not form S x
    // solver
    DBuf<double> E0, b, x, r, p, Ap, dx, rt, mud, partial, scal, tmp3n, fout;
    // precisions
    Prec<float> pf;
    Prec<double> pd;

    // last evaluation
    std::vector<double> last_pos, last_box;
    bool have_last = false;
    bool mu_valid = false;
    double energy_correction = 0.0;
    bool correction_pending = false;  // the solver's correction dot is on its way to hscal[2]

    // profiling: bit c of prof_mask times the work booked to category c
    unsigned prof_mask = 0;
    bool capturing = false;  // recording the solver's iteration graph: no timing events
    // the categories of the work inside one solver iteration; timing any of
    // them makes the iterations run eagerly instead of from the graph
    static constexpr unsigned kIterationCats = (1u << P_SPREAD) | (1u << P_FFT) | (1u << P_GATHER) |
                                               (1u << P_INDUCED) | (1u << P_COV) | (1u << P_SOLVER);
    struct Ev { int cat; hipEvent_t a, b; };
    std::vector<Ev> events;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    double prof_ms[MPID_PROF_COUNT] = {};
    int prof_count[MPID_PROF_COUNT] = {};

    ~mpid_context() {
        // Work still in flight must end before its streams, events, pinned
        // words and buffers go: a call that ended in an error can leave the
        // solver's issued-ahead iteration (which writes hscal through its
        // mapping) or the check stream running.
        for (hipStream_t s : {stream, rstream, vstream})
            if (s) (void)hipStreamSynchronize(s);
        for (auto e : pool) (void)hipEventDestroy(e);
        if (cg_exec) (void)hipGraphExecDestroy(cg_exec);
        if (hscal) (void)hipHostFree(hscal);
        if (hnl) (void)hipHostFree(hnl);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_sync) (void)hipEventDestroy(ev_sync);
        for (hipEvent_t e : ev_check)
            if (e) (void)hipEventDestroy(e);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (rstream && rstream != stream) (void)hipStreamDestroy(rstream);
        if (vstream && vstream != stream) (void)hipStreamDestroy(vstream);
        if (ev_v) (void)hipEventDestroy(ev_v);
        if (ev_vdone) (void)hipEventDestroy(ev_vdone);
        if (stream) (void)hipStreamDestroy(stream);
    }

    hipEvent_t get_event() {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[pool_used++];
    }
    int prof_override = -1;  // when set, every timed region is booked to this category
    // device time of the work that f issues to stream s, booked to cat
    template <class F> void timed_on(hipStream_t s, int cat, F&& f) {
        if (prof_override >= 0) cat = prof_override;
        if (capturing || !((prof_mask >> cat) & 1u)) {
            f();
            return;
        }
        hipEvent_t a = get_event(), b = get_event();
        HIP_CHECK(hipEventRecord(a, s));
        f();
        HIP_CHECK(hipEventRecord(b, s));
        events.push_back({cat, a, b});
    }
    template <class F> void timed(int cat, F&& f) { timed_on(stream, cat, std::forward<F>(f)); }
    void fork() {
        if (rstream == stream) return;
        HIP_CHECK(hipEventRecord(ev_fork, stream));
        HIP_CHECK(hipStreamWaitEvent(rstream, ev_fork, 0));
    }
    void join() {
        if (rstream == stream) return;
        HIP_CHECK(hipEventRecord(ev_join, rstream));
        HIP_CHECK(hipStreamWaitEvent(stream, ev_join, 0));
    }
    // vstream waits for the work issued to stream so far
    void mark_v() {
        if (vstream == stream) return;
        HIP_CHECK(hipEventRecord(ev_v, stream));
        HIP_CHECK(hipStreamWaitEvent(vstream, ev_v, 0));
    }
    // Runs f, which issues work to stream and rstream, on vstream instead (in
    // order there, after the last mark_v()).
    template <class F> void on_vstream(F&& f) {
        const hipStream_t s0 = stream, r0 = rstream;
        stream = rstream = vstream;
        try {
            f();
        } catch (...) {
            stream = s0;
            rstream = r0;
            throw;
        }
        stream = s0;
        rstream = r0;
    }
    // stream waits for the work issued to vstream
    void join_v() {
        if (vstream == stream) return;
        HIP_CHECK(hipEventRecord(ev_vdone, vstream));
        HIP_CHECK(hipStreamWaitEvent(stream, ev_vdone, 0));
    }
    void collect_profile() {
        if (!prof_mask) return;
        HIP_CHECK(hipStreamSynchronize(vstream));
        HIP_CHECK(hipStreamSynchronize(rstream));
        HIP_CHECK(hipStreamSynchronize(stream));
        for (auto& e : events) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
            prof_ms[e.cat] += ms;
            prof_count[e.cat] += 1;
        }
        events.clear();
        pool_used = 0;
    }
    void check_launch() { HIP_CHECK(hipGetLastError()); }

    // ---- set-up --------------------------------------------------------------
    void set_parameters(const mpid_parameters* P, bool first) {
        if (!P || P->n <= 0) throw MpidError(MPID_ERR_INVALID, "MPIDForce: no particles");
        npol_cached = -1;  // the polarisable set may change
        if (!first && P->n != n)
            throw MpidError(MPID_ERR_INVALID, "updateParametersInContext: The number of multipoles has changed");
        n = P->n;
        if (!pme() && n > kNoCutoffMaxAtoms)
            throw MpidError(MPID_ERR_UNSUPPORTED, "NoCutoff is limited to " + std::to_string(kNoCutoffMaxAtoms) +
                                                      " particles on this platform (every pair is listed)");
        // validation, MPIDForceImpl.cpp:51-150
        for (int i = 0; i < n; i++) {
            const double* q = P->quadrupoles + 6 * i;
            const double* o = P->octopoles + 10 * i;
            double tr = std::fabs(q[0] + q[2] + q[5]);
            if (tr > 1e-5)
                throw MpidError(MPID_ERR_INVALID, "MPIDForce: quadrupole for particle=" + std::to_string(i) +
                                                      " has nonzero trace: " + std::to_string(tr) +
                                                      "; MPID plugin assumes traceless quadrupole.");
            if (std::fabs(o[0] + o[2] + o[7]) > 1e-5 || std::fabs(o[1] + o[3] + o[8]) > 1e-5 ||
                std::fabs(o[4] + o[6] + o[9]) > 1e-5)
                throw MpidError(MPID_ERR_INVALID, "MPIDForce: octopole for particle=" + std::to_string(i) +
                                                      " has nonzero trace; MPID plugin assumes traceless octopoles.");
            int t = P->axis_types[i];
            if (t < 0 || t > 5)
                throw MpidError(MPID_ERR_INVALID, "MPIDForce: axis type=" + std::to_string(t) + " not currently handled");
            if (t != 5 && (P->atom_z[i] < 0 || P->atom_z[i] >= n))
                throw MpidError(MPID_ERR_INVALID, "MPIDForce: invalid z axis particle: " + std::to_string(P->atom_z[i]));
            if (t != 5 && t != 4 && (P->atom_x[i] < 0 || P->atom_x[i] >= n))
                throw MpidError(MPID_ERR_INVALID, "MPIDForce: invalid x axis particle: " + std::to_string(P->atom_x[i]));
            if ((t == 2 || t == 3) && (P->atom_y[i] < 0 || P->atom_y[i] >= n))
                throw MpidError(MPID_ERR_INVALID, "MPIDForce: invalid y axis particle: " + std::to_string(P->atom_y[i]));
            for (int k = 0; k < 3; k++)
                if (!(P->polarity[3 * i + k] >= 0.0))
                    throw MpidError(MPID_ERR_INVALID, "MPIDForce: negative polarizability for particle " + std::to_string(i));
        }
        // molecular-frame records in the internal orderings
        std::vector<double> m((size_t)n * kMP);
        std::vector<int4> ax(n);
        for (int i = 0; i < n; i++) {
            double* r = &m[(size_t)i * kMP];
            const double* q = P->quadrupoles + 6 * i;
            const double* o = P->octopoles + 10 * i;
            r[0] = P->charges[i];
            r[1] = P->dipoles[3 * i];
            r[2] = P->dipoles[3 * i + 1];
            r[3] = P->dipoles[3 * i + 2];
            // XX XY YY XZ YZ ZZ -> XX XY XZ YY YZ ZZ  (MPIDReferenceForce.cpp:305-310)
            r[4] = q[0]; r[5] = q[1]; r[6] = q[3]; r[7] = q[2]; r[8] = q[4]; r[9] = q[5];
            // XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ -> XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ (:312-321)
            r[10] = o[0]; r[11] = o[1]; r[12] = o[4]; r[13] = o[2]; r[14] = o[5];
            r[15] = o[7]; r[16] = o[3]; r[17] = o[6]; r[18] = o[8]; r[19] = o[9];
            int t = P->axis_types[i];
            ax[i] = make_int4(t == 5 ? -1 : P->atom_z[i], (t == 5 || t == 4) ? -1 : P->atom_x[i],
                              (t == 5 || t == 4) ? -1 : P->atom_y[i], t);
        }
        // covalent scale classes: the scale of the pair (i<j) comes from the
        // lists of i, later list types overwrite earlier ones; only classes
        // with a scale other than 1 are stored, for both atoms of the pair
        // (setupScaleMaps / getMultipoleScaleFactor, MPIDReferenceForce.cpp:190-241).
        // A 1-4 pair at scale14 = 1 is an ordinary pair (its Thole width is
        // the default one, as for every pair with a nonzero polarization
        // scale), so it is not stored either.  There is no limit on the
        // number of partners: the slots per atom follow the largest count.
        std::vector<std::vector<std::pair<int, int>>> plist(n);
        for (int a = 0; a < n; a++) {
            std::vector<std::pair<int, int>> mine;  // (partner, class) with partner > a
            for (int t = 0; t < 4; t++)
                for (int k = P->covalent_offsets[4 * a + t]; k < P->covalent_offsets[4 * a + t + 1]; k++) {
                    int bidx = P->covalent_atoms[k];
                    if (bidx < 0 || bidx >= n)
                        throw MpidError(MPID_ERR_INVALID, "MPIDForce: covalent atom index out of range");
                    if (bidx <= a) continue;
                    bool found = false;
                    for (auto& e : mine)
                        if (e.first == bidx) { e.second = t; found = true; }
                    if (!found) mine.push_back({bidx, t});
                }
            for (auto& e : mine) {
                int cls;
                if (e.second <= 1) cls = kScaleZero;            // 1-2, 1-3: scale 0
                else if (e.second == 2 && opt.scale14 != 1.0) cls = kScale14;  // 1-4: scale14
                else continue;                                  // 1-5 (and 1-4 at scale14 = 1): scale 1
                plist[a].push_back({e.first, cls});
                plist[e.first].push_back({a, cls});
            }
        }
        int stride = 1;
        for (int a = 0; a < n; a++) stride = std::max(stride, (int)plist[a].size());
        std::vector<int> part((size_t)n * stride, 0), np(n);
        for (int a = 0; a < n; a++) {
            np[a] = (int)plist[a].size();
            for (int k = 0; k < np[a]; k++) part[(size_t)a * stride + k] = (plist[a][k].first << 2) | plist[a][k].second;
        }
        pstride = stride;
        spart.resize((size_t)n * pstride);
        std::vector<double> th(P->tholes, P->tholes + n), po(P->polarity, P->polarity + 3 * n);
        mol.upload(m.data(), m.size(), stream);
        axes.upload(ax.data(), ax.size(), stream);
        thole.upload(th.data(), th.size(), stream);
        pol.upload(po.data(), po.size(), stream);
        partners.upload(part.data(), part.size(), stream);
        npart.upload(np.data(), np.size(), stream);
        HIP_CHECK(hipStreamSynchronize(stream));
        have_last = false;
        mu_valid = false;
    }

    // NoCutoff (the reference's non-PME MPIDReferenceForce): every pair, no
    // periodic images, no reciprocal or self terms.  It runs the PME path's
    // real-space kernels with alpha = 0 (their formulas then reduce to the
    // bare damped and scaled interactions), a cutoff that every pair passes,
    // a box of zero edges (no wrapping, no images) and no reciprocal cycles.
    // The lists then hold every atom, so systems are limited in size.
    static constexpr double kNoCutoffDistance = 1.0e6;  // nm
    static constexpr int kNoCutoffMaxAtoms = 16384;
    bool pme() const { return opt.nonbonded_method == MPID_NONBONDED_PME; }

    void set_options(const mpid_options* o) {
        if (!o) throw MpidError(MPID_ERR_INVALID, "missing options");
        opt = *o;
        if (opt.nonbonded_method != MPID_NONBONDED_PME && opt.nonbonded_method != MPID_NONBONDED_NOCUTOFF)
            throw MpidError(MPID_ERR_INVALID, "unknown nonbonded method");
        if (!pme()) {
            opt.ewald_alpha = 0.0;
            opt.cutoff = kNoCutoffDistance;
            opt.grid[0] = opt.grid[1] = opt.grid[2] = 0;
        }
        if (opt.polarization != MPID_POLARIZATION_MUTUAL && opt.polarization != MPID_POLARIZATION_DIRECT &&
            opt.polarization != MPID_POLARIZATION_EXTRAPOLATED)
            throw MpidError(MPID_ERR_INVALID, "unknown polarization type");
        {
            // coefficients copied here (the caller's array need not outlive the call)
            static const double kDefault[4] = {-0.154, 0.017, 0.658, 0.474};  // MPIDForce.cpp:46-49
            const int nc = opt.n_extrapolation > 0 ? opt.n_extrapolation : 4;
            const double* c = opt.n_extrapolation > 0 ? opt.extrapolation_coefficients : kDefault;
            if (opt.n_extrapolation < 0 || (opt.n_extrapolation > 0 && !c))
                throw MpidError(MPID_ERR_INVALID, "invalid extrapolation coefficients");
            if (nc > kMaxExtrapolation)
                throw MpidError(MPID_ERR_UNSUPPORTED, "more than " + std::to_string(kMaxExtrapolation) +
                                                          " extrapolation coefficients");
            ExtrapCoefficients cp{};
            cp.n = nc;
            for (int i = 0; i < nc; i++)
                for (int j = i; j < nc; j++) cp.c[i] += c[j];
            ext_cp = cp;
            opt.n_extrapolation = nc;
            opt.extrapolation_coefficients = nullptr;
        }
        if (pme() && !(opt.ewald_alpha > 0.0))
            throw MpidError(MPID_ERR_INVALID, "the Ewald alpha must be given explicitly (setPMEParameters)");
        if (!(opt.cutoff > 0.0)) throw MpidError(MPID_ERR_INVALID, "invalid cutoff");
        if (opt.precision != MPID_PRECISION_MIXED && opt.precision != MPID_PRECISION_DOUBLE)
            throw MpidError(MPID_ERR_INVALID, "invalid precision");
        if (opt.max_iterations < 0) throw MpidError(MPID_ERR_INVALID, "invalid iteration limit");
        if (opt.flags & ~(MPID_FLAG_CONSERVATIVE_TORQUES | MPID_FLAG_SINGLE_STREAM | MPID_FLAG_DIIS))
            throw MpidError(MPID_ERR_INVALID, "unknown option flags");
        if ((opt.flags & MPID_FLAG_DIIS) && opt.precision != MPID_PRECISION_DOUBLE)
            throw MpidError(MPID_ERR_UNSUPPORTED, "the DIIS solver runs in double precision only");
        for (int k = 0; k < 3 && pme(); k++) {
            int rr[16], c;
            FftAxis ax;
            if (opt.grid[k] < kOrder || opt.grid[k] > 1024 || !factor_radices(opt.grid[k], rr, c) ||
                !make_fft_axis(opt.grid[k], ax))
                throw MpidError(MPID_ERR_UNSUPPORTED, "PME grid dimension " + std::to_string(opt.grid[k]) +
                                                          " is not of the form 2^a 3^b 5^c between 6 and 1024");
        }
    }

    template <class R> void init_prec(Prec<R>& P) {
        if (P.ready) return;
        const size_t G = (size_t)opt.grid[0] * opt.grid[1] * opt.grid[2];
        P.pos.resize(n);
        P.lab.resize((size_t)n * kMP);
        P.td.resize(n);
        P.mu.resize(3 * (size_t)n);
        P.erec.resize(3 * (size_t)n);
        // the solver gathers the field at polarisable sites only; the other
        // entries stay zero (they are multiplied by S = 0, which needs them finite)
        HIP_CHECK(hipMemsetAsync(P.erec.p, 0, P.erec.n * sizeof(R), stream));
        P.phi.resize(35 * (size_t)n);
        P.phidp.resize(35 * (size_t)n);
        if (!pme()) {
            // no reciprocal space: the potentials and fields it would give stay zero
            HIP_CHECK(hipMemsetAsync(P.phi.p, 0, P.phi.n * sizeof(R), stream));
            HIP_CHECK(hipMemsetAsync(P.phidp.p, 0, P.phidp.n * sizeof(R), stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            P.ready = true;
            return;
        }
        P.grid.resize(G);
        P.spec.resize((size_t)opt.grid[0] * opt.grid[1] * (opt.grid[2] / 2 + 1));
        std::vector<double> mods;
        for (int k = 0; k < 3; k++) {
            std::vector<double> m = bspline_moduli(opt.grid[k]);
            for (double v : m) mods.push_back(v);
            const int nk = opt.grid[k];
            std::vector<R2<R>> tw(nk);
            for (int j = 0; j < nk; j++) {
                double ang = -2.0 * M_PI * j / nk;
                tw[j].x = (R)std::cos(ang);
                tw[j].y = (R)std::sin(ang);
            }
            P.tw[k].upload(tw.data(), tw.size(), stream);
            P.plan.tw[k] = P.tw[k].p;
            if (!make_fft_axis(nk, P.plan.ax[k])) throw MpidError(MPID_ERR_UNSUPPORTED, "PME grid size not supported");
        }
        P.moduli.upload(mods.data(), mods.size(), stream);
        P.infl.resize(P.spec.n);
        P.infl_valid = false;
        if (tiled_spread && !owner_spread)
            P.scratch.resize((size_t)tgeom.ntiles * tgeom.W[0] * tgeom.W[1] * tgeom.W[2]);
        HIP_CHECK(hipStreamSynchronize(stream));
        P.ready = true;
    }

    void alloc_state() {
        const size_t n3 = 3 * (size_t)n;
        pos_in.resize(n3);
        keys.resize(n);
        keys_sorted.resize(n);
        vals.resize(n);
        perm.resize(n);
        inv.resize(n);
        sort_tmp.resize(std::max<size_t>(sort_temp_bytes(n), 1));
        spos.resize(n);
        smol.resize((size_t)n * kMP);
        saxes.resize(n);
        sthole.resize(n);
        spol.resize(n3);
        snpart.resize(n);
        lab.resize((size_t)n * kMP);
        labpol.resize(6 * (size_t)n);
        sqrtpol.resize(6 * (size_t)n);
        damp.resize(n);
        aniso.resize(n);
        chg.resize(n);
        const int nb = (n + kTile - 1) / kTile;
        bmin.resize(nb);
        bmax.resize(nb);
        brow_count.resize(nb);
        brow_start.resize(nb + 1);
        brow_upper.resize(nb);
        brow_ustart.resize(nb + 1);
        ncount.resize(n);
        pcount.resize(n);
        polcnt.resize(nb);
        polmask.resize(nb);
        polstart.resize(nb + 1);
        pol_idx.resize(n);
        nl_scalars.resize(2);
        field_acc.resize(n3);
        force_acc.resize(n3);
        torque_acc.resize(n3);
        energy_acc.resize(2 * kEnergySlots);
        E0.resize(n3);
        b.resize(n3);
        x.resize(n3);
        r.resize(n3);
        p.resize(n3);
        Ap.resize(n3);
        dx.resize(n3);
        rt.resize(n3);
        mud.resize(n3);
        tmp3n.resize(n3);
        fout.resize(n3);
        partial.resize(2 * 512);
        scal.resize(kScalars);
        choose_tiles();
        if (opt.precision == MPID_PRECISION_MIXED) init_prec(pf);
        init_prec(pd);  // FP64 path (also the refinement residual of the mixed path)
    }

    // The spreading's tiling: opt.spread_level / spread_mode when given (one
    // of spread_tilings(opt.grid)), else the first of that list.
    void choose_tiles() {
        tiled_spread = false;
        owner_spread = false;
        if (!pme()) return;
        const std::vector<SpreadTiling> avail = spread_tilings(opt.grid);
        SpreadTiling pick = avail.front();
        if (opt.spread_level != 0) {
            bool found = false;
            for (const SpreadTiling& t : avail)
                if (t.level == opt.spread_level && t.mode == opt.spread_mode) found = true;
            if (!found)
                throw MpidError(MPID_ERR_INVALID, "spread tiling (level " + std::to_string(opt.spread_level) +
                                                      ", mode " + std::to_string(opt.spread_mode) +
                                                      ") is not available for this PME grid");
            pick = SpreadTiling{opt.spread_level, opt.spread_mode};
        }
        if (pick.level > 0) {
            tile_geometry(opt.grid, pick.level, tgeom);
            tiled_spread = true;
            owner_spread = pick.mode == kSpreadOwner;
        }
        if (tiled_spread) {
            tstart.resize(tgeom.ntiles);
            tend.resize(tgeom.ntiles);
            tpstart.resize(tgeom.ntiles);
            tpend.resize(tgeom.ntiles);
        }
    }

    // ---- evaluation ------------------------------------------------------------
    Box<double> dbox{};
    double box_vec[9];
    // box^-1 (the reciprocal box vectors as rows), lower triangular:
    // MPIDReferenceForce.cpp:2627-2632
    double rinv[3][3] = {};

    void set_box(const double* boxv) {
        if (!pme()) {
            // no periodicity: zero edges make every wrap and image shift vanish
            for (int k = 0; k < 9; k++) box_vec[k] = boxv ? boxv[k] : 0.0;
            dbox = Box<double>{};
            for (auto& row : rinv)
                for (double& v : row) v = 0.0;
            return;
        }
        if (!boxv) throw MpidError(MPID_ERR_INVALID, "PME requires periodic box vectors");
        // Box vectors in OpenMM's reduced form: a = (ax, 0, 0), b = (bx, by, 0),
        // c = (cx, cy, cz) with |bx|, |cx| <= ax / 2 and |cy| <= by / 2
        // (the form the reference's getPeriodicDelta and reciprocal box assume)
        if (boxv[1] != 0 || boxv[2] != 0 || boxv[5] != 0)
            throw MpidError(MPID_ERR_INVALID,
                            "box vectors must be in reduced form: a along x, b in the xy plane");
        const double tol = 1.0 + 1e-9;
        if (std::fabs(boxv[3]) > 0.5 * boxv[0] * tol || std::fabs(boxv[6]) > 0.5 * boxv[0] * tol ||
            std::fabs(boxv[7]) > 0.5 * boxv[4] * tol)
            throw MpidError(MPID_ERR_INVALID,
                            "box vectors must be in reduced form: |bx|, |cx| <= ax / 2 and |cy| <= by / 2");
        for (int k = 0; k < 9; k++) box_vec[k] = boxv[k];
        for (int k = 0; k < 3; k++) {
            double L = boxv[4 * k];
            if (!(L > 0.0)) throw MpidError(MPID_ERR_INVALID, "Box size of zero is invalid.");
            // MPIDReferenceKernels.cpp:193-197 and MPIDForceImpl.cpp:61-66
            if (L < 1.999999 * opt.cutoff)
                throw MpidError(MPID_ERR_INVALID,
                                "The periodic box size has decreased to less than twice the nonbonded cutoff.");
            dbox.L[k] = L;
            dbox.inv[k] = 1.0 / L;
        }
        dbox.bx = boxv[3];
        dbox.cx = boxv[6];
        dbox.cy = boxv[7];
        const double ax = boxv[0], by = boxv[4], cz = boxv[8];
        const double sc = 1.0 / (ax * by * cz);
        for (auto& row : rinv)
            for (double& v : row) v = 0.0;
        rinv[0][0] = by * cz * sc;
        rinv[1][0] = -dbox.bx * cz * sc;
        rinv[1][1] = ax * cz * sc;
        rinv[2][0] = (dbox.bx * dbox.cy - by * dbox.cx) * sc;
        rinv[2][1] = -ax * dbox.cy * sc;
        rinv[2][2] = ax * by * sc;
    }
    bool triclinic() const { return dbox.bx != 0.0 || dbox.cx != 0.0 || dbox.cy != 0.0; }
    template <class R> Box<R> rbox() const {
        Box<R> bx;
        for (int k = 0; k < 3; k++) {
            bx.L[k] = (R)dbox.L[k];
            bx.inv[k] = (R)dbox.inv[k];
        }
        bx.bx = (R)dbox.bx;
        bx.cx = (R)dbox.cx;
        bx.cy = (R)dbox.cy;
        return bx;
    }
    // Grid-coordinate transform A_jk = n_k (box^-1)_jk
    // (MPIDReferenceForce.cpp:3083 and 3177)
    template <class R> FracT<R> frac_transform() const {
        FracT<R> A;
        A.a00 = (R)(opt.grid[0] * rinv[0][0]);
        A.a10 = (R)(opt.grid[0] * rinv[1][0]);
        A.a11 = (R)(opt.grid[1] * rinv[1][1]);
        A.a20 = (R)(opt.grid[0] * rinv[2][0]);
        A.a21 = (R)(opt.grid[1] * rinv[2][1]);
        A.a22 = (R)(opt.grid[2] * rinv[2][2]);
        A.tri = triclinic() ? 1 : 0;
        return A;
    }
    template <class R> PmeGeom<R> geom() const {
        PmeGeom<R> g;
        for (int k = 0; k < 3; k++) {
            g.n[k] = opt.grid[k];
            g.gscale[k] = (R)(opt.grid[k] * dbox.inv[k]);
        }
        g.A = frac_transform<R>();
        g.box = rbox<R>();
        g.skip = iteration_skip();
        return g;
    }
    // Kernels recorded into the solver's iteration graph take the stop flag:
    // an iteration issued after the solve has stopped does no work.
    const double* iteration_skip() const { return capturing ? scal.p + kCgStop : nullptr; }
    template <class R> PairParams<R> pair_params() const {
        PairParams<R> P;
        P.box = rbox<R>();
        P.n = n;
        P.rc2 = (R)(opt.cutoff * opt.cutoff);
        P.alpha = (R)opt.ewald_alpha;
        // (2 alpha^2)^k / (alpha sqrt(pi)); all zero at alpha = 0 (NoCutoff)
        double a = opt.ewald_alpha, aln = a > 0.0 ? 1.0 / (kSqrtPi * a) : 0.0;
        P.aln[0] = (R)0;
        for (int k = 1; k <= 4; k++) {
            aln *= 2.0 * a * a;
            P.aln[k] = (R)aln;
        }
        P.default_thole = (R)opt.default_thole;
        P.scale14 = (R)opt.scale14;
        P.mutual = opt.polarization == MPID_POLARIZATION_MUTUAL ? 1 : 0;
        P.fix_torque = (opt.flags & MPID_FLAG_CONSERVATIVE_TORQUES) ? 1 : 0;
        P.skip = iteration_skip();
        return P;
    }
    template <class R> PairAtoms<R> pair_atoms(Prec<R>& P) const {
        PairAtoms<R> A;
        A.pos = P.pos.p;
        A.lab = P.lab.p;
        A.td = P.td.p;
        A.aniso = aniso.p;
        A.partners = spart.p;
        A.npart = snpart.p;
        A.pstride = pstride;
        return A;
    }
    double selfc() const { return (4.0 / 3.0) * opt.ewald_alpha * opt.ewald_alpha * opt.ewald_alpha / kSqrtPi; }

    // positions (device, original order) -> sorted lab-frame state
    // all: build the lists of every atom (the inspection entry points, whose
    // field passes visit all atoms) instead of those of the polarisable atoms.
    // fixed_recip: issue the FP64 reciprocal cycle of the permanent multipoles
    // (to rstream, as soon as the lab-frame multipoles exist, so that it runs
    // beside the neighbour-list build); fixed_field() joins it.
    void prepare(const double* d_pos, bool all = false, bool fixed_recip = true) {
        all_lists = all;
        update_influence(pd);
        if (opt.precision == MPID_PRECISION_MIXED) update_influence(pf);
        timed(P_SORT, [&] {
            launch_morton_keys(stream, n, d_pos, dbox, keys.p, vals.p);
            int rc = launch_sort_pairs(stream, sort_tmp.p, sort_tmp.n, keys.p, keys_sorted.p, vals.p, perm.p, n);
            if (rc != 0) throw HipError("rocprim radix sort failed");
            launch_invert_perm(stream, n, perm.p, inv.p);
            launch_gather_sorted(stream, n, perm.p, inv.p, d_pos, dbox, mol.p, axes.p, thole.p, pol.p, partners.p,
                                 npart.p, pstride, spos.p, smol.p, saxes.p, sthole.p, spol.p, spart.p, snpart.p);
        });
        if (tiled_spread) {
            timed(P_SORT, [&] {
                HIP_CHECK(hipMemsetAsync(tstart.p, 0, tstart.n * sizeof(int), stream));
                HIP_CHECK(hipMemsetAsync(tend.p, 0, tend.n * sizeof(int), stream));
                launch_tile_ranges(stream, n, keys_sorted.p, nullptr, nullptr, 3 * (10 - tgeom.m), tstart.p, tend.p);
            });
        }
        timed(P_ROTATE, [&] {
            launch_rotate(stream, n, dbox, spos.p, saxes.p, smol.p, spol.p, lab.p, labpol.p, sqrtpol.p, damp.p,
                          aniso.p, chg.p);
            if (opt.precision == MPID_PRECISION_MIXED)
                launch_convert<float>(stream, n, spos.p, lab.p, sthole.p, damp.p, pf.pos.p, pf.lab.p, pf.td.p);
            launch_convert<double>(stream, n, spos.p, lab.p, sthole.p, damp.p, pd.pos.p, pd.lab.p, pd.td.p);
        });
        check_launch();
        // blocks, block rows, neighbour lists and tile masks.  The row count
        // is issued before the FP64 reciprocal cycle of the permanent
        // multipoles, so that the host issues that cycle while the device
        // counts, and it is read back only then.
        const int nb = (n + kTile - 1) / kTile;
        timed(P_TILES, [&] {
            launch_block_bounds<double>(stream, n, pd.pos.p, bmin.p, bmax.p);
            launch_block_rows(stream, nb, bmin.p, bmax.p, rbox<float>(), (float)opt.cutoff, nullptr, brow_count.p,
                              brow_upper.p, nullptr, nullptr, 0);
            launch_exclusive_scan(stream, nb, brow_count.p, brow_start.p);
            launch_exclusive_scan(stream, nb, brow_upper.p, brow_ustart.p);
        });
        check_launch();
        // After a verified build for the current parameters the count is
        // not waited for: the row kernels run over the capacity and stop at
        // the count on the device, which is checked with the lists at the
        // solver's first synchronisation (check_lists_pending).
        rows_pending = !all && npol_cached >= 0 && brows.n > 0;
        if (rows_pending) {
            HIP_CHECK(hipMemcpyAsync(hnl + 2, brow_start.p + nb, sizeof(int), hipMemcpyDeviceToHost, stream));
            if (fixed_recip) recip_cycle_fixed(pd);
        } else {
            HIP_CHECK(hipMemcpyAsync(hscal + 3, brow_start.p + nb, sizeof(int), hipMemcpyDeviceToHost, stream));
            if (fixed_recip) recip_cycle_fixed(pd);
            wait_stream();
            std::memcpy(&rows_total, hscal + 3, sizeof(int));
            grow_rows(rows_total);
            stats.ntiles = (rows_total + nb) / 2;  // block pairs with J >= I
        }
        build_neighbour_lists(nb);
        recip_cycle_fixed_rest(pd);  // if the list build has not issued it
    }
    // an upper bound of the electrostatics kernels' tile count: the row
    // entries with J >= I, of the verified count or of the capacity while
    // the count is pending
    int tile_bound() const {
        const int nb = (n + kTile - 1) / kTile;
        return ((rows_pending ? (int)brows.n : rows_total) + nb) / 2 + 1;
    }
    void grow_rows(int count) {
        if ((size_t)count <= brows.n) return;
        const size_t cap = count + count / 4 + 64;
        brows.resize(cap);
        brow_owner.resize(cap);
        tile_act.resize(cap);
        tmask.resize(cap * kTile);
        tmaskcc.resize(cap * kTile);
        nmask.resize(cap * kTile);
    }

    // Per-atom lists: of the polarisable atoms (by position in pol_idx) for an
    // evaluation, of every atom (by atom index) when all_lists is set.  The
    // capacity starts from the mean number of atoms in a cutoff sphere; a
    // list that overflows is rebuilt with a larger one.
    void build_neighbour_lists(int nb) {
        if (nl_cap == 0) {
            if (!pme()) {
                nl_cap = n;  // every atom is a neighbour
            } else {
                const double vol = dbox.L[0] * dbox.L[1] * dbox.L[2];
                const double mean = n / vol * (4.0 / 3.0) * M_PI * opt.cutoff * opt.cutoff * opt.cutoff;
                nl_cap = std::min<int>((int)(1.3 * mean) + 64, n);
            }
        }
        timed(P_TILES, [&] {
            launch_block_rows(stream, nb, bmin.p, bmax.p, rbox<float>(), (float)opt.cutoff, brow_start.p, nullptr,
                              nullptr, brows.p, brow_owner.p, (int)brows.n);
        });
        for (int attempt = 0; attempt < 2; attempt++) {
            nl_cap = (nl_cap + kNlSplit - 1) / kNlSplit * kNlSplit;
            nbr.resize((size_t)n * nl_cap);
            pnbr.resize((size_t)n * nl_cap);
            timed(P_TILES, [&] {
                HIP_CHECK(hipMemsetAsync(nl_scalars.p, 0, 2 * sizeof(int), stream));
                launch_nl_build(stream, n, rows_pending ? (int)brows.n : rows_total, (int)brows.n, all_lists,
                                brow_start.p, brows.p, brow_owner.p, bmin.p, bmax.p,
                                pd.pos.p, spart.p, snpart.p, pstride, dbox, opt.cutoff * opt.cutoff, nl_cap, nmask.p,
                                polmask.p, polcnt.p, polstart.p, nbr.p, ncount.p, nl_scalars.p, tmask.p, pnbr.p,
                                pcount.p, chg.p, tmaskcc.p, tile_act.p, brow_ustart.p);
                launch_pol_fill(stream, n, pd.pos.p, polstart.p, pol_idx.p);
                if (tiled_spread) {
                    HIP_CHECK(hipMemsetAsync(tpstart.p, 0, tpstart.n * sizeof(int), stream));
                    HIP_CHECK(hipMemsetAsync(tpend.p, 0, tpend.n * sizeof(int), stream));
                    launch_tile_ranges(stream, n, keys_sorted.p, pol_idx.p, polstart.p + nb, 3 * (10 - tgeom.m),
                                       tpstart.p, tpend.p);
                }
            });
            check_launch();
            recip_cycle_fixed_rest(pd);
            if (!all_lists && npol_cached >= 0) {
                HIP_CHECK(hipMemcpyAsync(hnl, nl_scalars.p, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
                npol = npol_cached;
                lists_pending = true;
                return;
            }
            int scal[2];
            HIP_CHECK(hipMemcpyAsync(scal, nl_scalars.p, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipMemcpyAsync(&npol, polstart.p + nb, sizeof(int), hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            if (scal[1] != 0) throw MpidError(MPID_ERR_HIP, "neighbour masks: a block pair lacks its mirror entry");
            nl_max = scal[0];
            if (nl_max <= nl_cap) {
                if (!all_lists) npol_cached = npol;
                return;
            }
            nl_cap = nl_max + nl_max / 8 + kNlSplit;
        }
        throw MpidError(MPID_ERR_HIP, "neighbour list overflow after regrowth");
    }
    // The checks of an unverified build (after a synchronisation that covers
    // it).  On overflow the capacity grows and ListsOverflow asks run() to
    // evaluate again; no output has been written by then.
    struct ListsOverflow {};
    void check_lists_pending() {
        if (!lists_pending) return;
        lists_pending = false;
        if (rows_pending) {
            rows_pending = false;
            const int count = hnl[2];
            if ((size_t)count > brows.n) {  // rows beyond the capacity were dropped: grow and evaluate again
                rows_grow_to = count;
                npol_cached = -1;
                throw ListsOverflow{};
            }
            rows_total = count;
            stats.ntiles = (rows_total + (n + kTile - 1) / kTile) / 2;
        }
        if (hnl[1] != 0) throw MpidError(MPID_ERR_HIP, "neighbour masks: a block pair lacks its mirror entry");
        nl_max = hnl[0];
        if (nl_max > nl_cap) {
            nl_cap = nl_max + nl_max / 8 + kNlSplit;
            npol_cached = -1;
            throw ListsOverflow{};
        }
    }
    NbrList nlist() const { return NbrList{nbr.p, ncount.p}; }
    NbrList nlist_pol() const { return NbrList{pnbr.p, pcount.p}; }

    template <class R> ConvParams<R> conv_params(const Prec<R>& P) const {
        ConvParams<R> c;
        c.infl = P.infl.p;
        c.skip = iteration_skip();
        return c;
    }
    // The influence-function table of a precision, recomputed (on stream)
    // when the box or alpha differs from the ones it holds.
    template <class R> void update_influence(Prec<R>& P) {
        if (!pme() || !P.ready) return;
        double key[10];
        for (int k = 0; k < 9; k++) key[k] = box_vec[k];
        key[9] = opt.ewald_alpha;
        if (P.infl_valid && std::memcmp(key, P.infl_key, sizeof(key)) == 0) return;
        InflParams ip;
        for (int k = 0; k < 3; k++) ip.n[k] = opt.grid[k];
        ip.expf = M_PI * M_PI / (opt.ewald_alpha * opt.ewald_alpha);
        ip.sc = 1.0 / (M_PI * dbox.L[0] * dbox.L[1] * dbox.L[2]);
        ip.r00 = rinv[0][0];
        ip.r10 = rinv[1][0];
        ip.r11 = rinv[1][1];
        ip.r20 = rinv[2][0];
        ip.r21 = rinv[2][1];
        ip.r22 = rinv[2][2];
        timed(P_FFT, [&] { launch_influence_table<R>(stream, ip, P.moduli.p, P.infl.p); });
        check_launch();
        std::memcpy(P.infl_key, key, sizeof(key));
        P.infl_valid = true;
    }

    // Reciprocal cycles: issued to rstream after a fork(); the caller joins
    // before it reads their output (phi, phidp, erec).
    // Issued in two parts: the spreading first, the transforms and the gather
    // once the neighbour-list kernels are issued (the host's launches would
    // otherwise hold those back while the spreading runs).
    bool fixed_recip_rest = false;
    template <class R> void recip_cycle_fixed(Prec<R>& P) {
        if (!pme()) return;  // phi stays zero
        PmeGeom<R> g = geom<R>();
        fork();
        const size_t G = P.grid.n;
        timed_on(rstream, P_SPREAD, [&] {
            if (owner_spread) {
                launch_spread_fixed_owner<R>(rstream, P.pos.p, P.lab.p, g, tgeom, tstart.p, tend.p, P.grid.p);
                return;
            }
            launch_fill<R>(rstream, P.grid.p, G, R(0));
            if (tiled_spread) {
                launch_spread_fixed_tiled<R>(rstream, P.pos.p, P.lab.p, g, tgeom, tstart.p, tend.p, P.scratch.p,
                                             P.grid.p);
                if (fold_by_tiles(tgeom)) launch_fold_tiles<R>(rstream, g, tgeom, P.scratch.p, P.grid.p);
                else launch_fold<R>(rstream, g, tgeom, P.scratch.p, P.grid.p);
            } else {
                launch_spread_fixed<R>(rstream, n, P.pos.p, P.lab.p, g, P.grid.p);
            }
        });
        check_launch();
        fixed_recip_rest = true;
    }
    template <class R> void recip_cycle_fixed_rest(Prec<R>& P) {
        if (!fixed_recip_rest) return;
        fixed_recip_rest = false;
        PmeGeom<R> g = geom<R>();
        timed_on(rstream, P_FFT, [&] { launch_recip_convolution<R>(rstream, P.plan, P.grid.p, P.spec.p, conv_params(P)); });
        timed_on(rstream, P_GATHER, [&] { launch_gather_phi<R>(rstream, n, P.pos.p, g, P.grid.p, P.phi.p); });
        check_launch();
    }

    // Reciprocal part of the field of the induced dipoles in P.mu.  Unless
    // all_atoms, mu is known to vanish on unpolarisable sites: only the
    // polarisable ones are spread, and (full == false) only their field is
    // gathered.  full: the 35 potential derivatives at every atom (forces).
    // forked: the caller has already made rstream wait for its work
    template <class R> void recip_cycle_induced(Prec<R>& P, bool full, bool all_atoms = false, bool forked = false) {
        if (!pme()) return;  // erec and phidp stay zero
        PmeGeom<R> g = geom<R>();
        if (!forked) fork();
        const size_t G = P.grid.n;
        timed_on(rstream, P_SPREAD, [&] {
            if (owner_spread) {
                if (all_atoms)
                    launch_spread_induced_owner<R>(rstream, P.pos.p, P.mu.p, g, tgeom, nullptr, tstart.p, tend.p,
                                                   P.grid.p);
                else
                    launch_spread_induced_owner<R>(rstream, P.pos.p, P.mu.p, g, tgeom, pol_idx.p, tpstart.p,
                                                   tpend.p, P.grid.p);
                return;
            }
            launch_fill<R>(rstream, P.grid.p, G, R(0));
            if (tiled_spread) {
                if (all_atoms)
                    launch_spread_induced_tiled<R>(rstream, P.pos.p, P.mu.p, g, tgeom, nullptr, tstart.p, tend.p,
                                                   P.scratch.p, P.grid.p);
                else
                    launch_spread_induced_tiled<R>(rstream, P.pos.p, P.mu.p, g, tgeom, pol_idx.p, tpstart.p, tpend.p,
                                                   P.scratch.p, P.grid.p);
                if (fold_by_tiles(tgeom)) launch_fold_tiles<R>(rstream, g, tgeom, P.scratch.p, P.grid.p);
                else launch_fold<R>(rstream, g, tgeom, P.scratch.p, P.grid.p);
            } else {
                launch_spread_induced<R>(rstream, n, P.pos.p, P.mu.p, g, P.grid.p);
            }
        });
        timed_on(rstream, P_FFT, [&] { launch_recip_convolution<R>(rstream, P.plan, P.grid.p, P.spec.p, conv_params(P)); });
        timed_on(rstream, P_GATHER, [&] {
            if (full) launch_gather_phi<R>(rstream, n, P.pos.p, g, P.grid.p, P.phidp.p);
            else if (all_atoms) launch_gather_field<R>(rstream, n, nullptr, P.pos.p, g, P.grid.p, P.erec.p);
            else launch_gather_field<R>(rstream, npol, pol_idx.p, P.pos.p, g, P.grid.p, P.erec.p);
        });
        check_launch();
    }

    // The permanent multipoles go through the reciprocal sum once per
    // evaluation, always in FP64: each atom's potential contains its own
    // smeared moments, whose energy nearly cancels the self energy, so FP32
    // rounding of that potential would dominate the error of the total.
    // Permanent field.  The solver needs it at the polarisable sites only;
    // all_atoms computes it everywhere (the inspection entry point).
    // the lists of the last build must be the ones the field pass reads
    void check_lists(bool all_atoms) const {
        if (all_atoms != all_lists)
            throw MpidError(MPID_ERR_HIP, "internal: neighbour lists built for another set of atoms");
    }
    template <class R> void fixed_field(Prec<R>& P, bool all_atoms = false) {
        check_lists(all_atoms);
        // the FP64 reciprocal cycle of the permanent multipoles was issued by
        // prepare() and runs beside the neighbour lists and these kernels
        timed(P_FIXED, [&] {
            launch_zero_acc(stream, field_acc.p, field_acc.n);
            launch_nl_fixed_field<R>(stream, all_atoms ? n : npol, all_atoms ? nullptr : pol_idx.p, nlist(),
                                     pair_atoms(P), pair_params<R>(), field_acc.p);
        });
        timed(P_COV, [&] {
            launch_cov_fixed_field<double>(stream, n, pair_atoms(pd), pair_params<double>(), field_acc.p);
        });
        join();
        timed(P_OTHER, [&] {
            launch_fixed_field_finish<double>(stream, n, pd.phi.p, lab.p, field_acc.p, frac_transform<double>(), selfc(), E0.p);
        });
        check_launch();
    }

    // Field T*mu of the induced dipoles in P.mu: accumulator + P.erec.  In the
    // solver (all_atoms false) mu vanishes on unpolarisable sites and the field
    // is only used on polarisable ones, so only pairs of polarisable sites are
    // visited; all_atoms gives the field of an arbitrary mu everywhere.
    // The reciprocal part goes to rstream first and runs beside the pair
    // kernels; the caller joins before it reads erec.
    template <class R> void induced_field(Prec<R>& P, bool all_atoms = false) {
        check_lists(all_atoms);
        recip_cycle_induced(P, false, all_atoms);
        timed(P_INDUCED, [&] {
            launch_zero_acc(stream, field_acc.p, field_acc.n, iteration_skip());
            if (all_atoms)
                launch_nl_induced_field<R>(stream, n, nullptr, nlist(), pair_atoms(P), P.mu.p, pair_params<R>(),
                                           field_acc.p);
            else
                launch_nl_induced_field<R>(stream, npol, pol_idx.p, nlist_pol(), pair_atoms(P), P.mu.p,
                                           pair_params<R>(), field_acc.p);
        });
        timed(P_COV, [&] {
            launch_cov_induced_field<R>(stream, n, pair_atoms(P), P.mu.p, pair_params<R>(), all_atoms ? 0 : 1,
                                        field_acc.p);
        });
    }

    // Av = v - S T S v
    template <class R> void matvec(Prec<R>& P, const double* v, double* Av) {
        timed(P_SOLVER, [&] { launch_apply_S<R>(stream, n, sqrtpol.p, v, P.mu.p, nullptr); });
        induced_field(P);
        join();
        timed(P_SOLVER, [&] {
            launch_matvec_finish<R>(stream, n, sqrtpol.p, field_acc.p, P.erec.p, P.mu.p, selfc(), v, Av);
        });
        check_launch();
        stats.matvecs++;
    }

    // Records f (one CG iteration on stream and rstream, which leaves the
    // residual norm in hscal[0]) into cg_exec: the executable graph is
    // updated in place when its topology is unchanged, else instantiated.
    template <class F> void capture_cg(F&& f) {
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        capturing = true;
        try {
            f();
        } catch (...) {
            capturing = false;
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(stream, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            throw;
        }
        capturing = false;
        HIP_CHECK(hipStreamEndCapture(stream, &g));
        bool updated = false;
        if (cg_exec) {
            hipGraphNode_t err = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(cg_exec, g, &err, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(cg_exec);
                cg_exec = nullptr;
            }
        }
        hipError_t e = hipSuccess;
        if (!updated) e = hipGraphInstantiate(&cg_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            cg_exec = nullptr;
            throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }

    // Waits for the work issued to stream so far by polling an event: the
    // solver's host decisions sit on this path, and polling resumes sooner
    // than a blocking synchronisation.
    void wait_stream() {
        HIP_CHECK(hipEventRecord(ev_sync, stream));
        wait_event(ev_sync);
    }
    void wait_event(hipEvent_t e) {
        for (;;) {
            const hipError_t q = hipEventQuery(e);
            if (q == hipSuccess) return;
            if (q != hipErrorNotReady) HIP_CHECK(q);
        }
    }
    double read_scalar(const double* d) {
        HIP_CHECK(hipMemcpyAsync(hscal + 1, d, sizeof(double), hipMemcpyDeviceToHost, stream));
        wait_stream();
        return hscal[1];
    }
    double debye_eps(double ss) const { return kDebye * std::sqrt(ss / n); }

    // Conjugate gradients on (I - S T S) sol = rhs.  With zero_start the
    // iteration starts from sol = 0 (r = rhs), else from the current sol.
    // Stops when the RMS-Debye norm of S r falls below stop_eps; returns it.
    //
    // The device makes the stopping test (check c after c iterations, see
    // launch_cg_begin) and the host reads it from pinned memory.  Replayed
    // from the graph, the next iteration is issued before the host reads the
    // previous check, so the device never waits for the host; an iteration
    // issued after the solve has stopped does no work.
    template <class R>
    double cg(Prec<R>& P, double* sol, const double* rhs, bool zero_start, double stop_eps, int& iters, int max_iter) {
        const int n3 = 3 * n;
        auto body = [&] {
            induced_field(P);
            join();
            timed(P_SOLVER, [&] {
                launch_cg_iteration<R>(stream, n, sqrtpol.p, field_acc.p, P.erec.p, P.mu.p, selfc(), p.p, Ap.p, sol,
                                       r.p, partial.p, scal.p, hscal_dev + kCheck, stop_eps);
            });
        };
        // The iteration graph is recorded first: the host does it while the
        // device still runs the work issued before.  (Timing events cannot sit
        // in a graph: when the iteration's categories are timed, the loop runs
        // eagerly and the host reads every check before the next iteration.)
        const bool eager = (prof_mask & kIterationCats) != 0;
        if (!eager && iters < max_iter) capture_cg(body);
        timed(P_SOLVER, [&] {
            if (zero_start) {
                launch_fill<double>(stream, sol, n3, 0.0);
                HIP_CHECK(hipMemcpyAsync(r.p, rhs, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream));
            }
        });
        if (!zero_start) {
            matvec(P, sol, Ap.p);
            timed(P_SOLVER, [&] { launch_sub(stream, n3, rhs, Ap.p, r.p); });
        }
        timed(P_SOLVER, [&] {
            HIP_CHECK(hipMemcpyAsync(p.p, r.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream));
            launch_residual_norm(stream, n, sqrtpol.p, r.p, partial.p, scal.p + 2);  // scal[2] = rr, scal[3] = ss
            HIP_CHECK(hipMemcpyAsync(scal.p, scal.p + 2, sizeof(double), hipMemcpyDeviceToDevice, stream));
            launch_cg_begin(stream, n, scal.p, hscal_dev + kCheck, stop_eps);
            // the first direction's operator input (each iteration leaves the
            // next one's behind); issued before the norm is known
            launch_apply_S<R>(stream, n, sqrtpol.p, p.p, P.mu.p, nullptr);
        });
        HIP_CHECK(hipEventRecord(ev_check[0], stream));
        const int iters0 = iters;
        int issued = 0, seen = 0;  // iterations issued; the check the host reads next
        auto issue = [&] {
            if (iters0 + issued >= max_iter) return;
            if (eager) {
                body();
                check_launch();
            } else {
                HIP_CHECK(hipGraphLaunch(cg_exec, stream));
            }
            issued++;
            HIP_CHECK(hipEventRecord(ev_check[issued & 1], stream));
        };
        if (!eager) issue();
        double eps;
        for (;;) {
            wait_event(ev_check[seen & 1]);
            if (seen == 0) check_lists_pending();
            const double* c = hscal + kCheck + 2 * (seen & 1);
            eps = debye_eps(c[0]);
            if (c[1] != 0.0) break;  // converged: an iteration issued after this one does nothing
            if (eager) issue();
            if (issued == seen) break;  // the iteration limit
            seen++;                     // iteration `seen` runs (issued, and the solve had not stopped)
            if (!eager) issue();        // one iteration ahead of the host
        }
        iters = iters0 + seen;
        stats.matvecs += seen;
        return eps;
    }

    // Extrapolated (OPT) polarization, convergeInduceDipolesByExtrapolation
    // (MPIDReferenceForce.cpp:1125-1180): mu_(0) = alpha E0 = S b, then for
    // m = 0 .. N-2 the field and field gradient of mu_(m) (the operator's
    // kernels with the gradient terms, the reciprocal potential derivatives
    // at every atom, the self term) and mu_(m+1) = alpha E_(m); finally
    // mu = sum_m c'_m mu_(m).  E_(m) and G_(m) stay for the response forces
    // of the final pass.
    template <class R> void extrapolate(Prec<R>& P) {
        check_lists(false);
        const int N = ext_cp.n;
        const size_t n3 = 3 * (size_t)n, n6 = 6 * (size_t)n;
        ptmu.resize((size_t)N * n3);
        ptE.resize((size_t)std::max(1, N - 1) * n3);
        ptG.resize((size_t)std::max(1, N - 1) * n6);
        grad_acc.resize(n6);
        timed(P_SOLVER, [&] { launch_apply_S<R>(stream, n, sqrtpol.p, b.p, P.mu.p, ptmu.p); });
        for (int m = 0; m + 1 < N; m++) {
            recip_cycle_induced(P, true);  // phidp of mu_(m) at every atom, on rstream
            timed(P_INDUCED, [&] {
                launch_zero_acc(stream, field_acc.p, field_acc.n);
                launch_zero_acc(stream, grad_acc.p, grad_acc.n);
                launch_nl_induced_field<R>(stream, npol, pol_idx.p, nlist_pol(), pair_atoms(P), P.mu.p,
                                           pair_params<R>(), field_acc.p, grad_acc.p);
            });
            timed(P_COV, [&] {
                launch_cov_induced_field<R>(stream, n, pair_atoms(P), P.mu.p, pair_params<R>(), 1, field_acc.p,
                                            grad_acc.p);
            });
            join();
            timed(P_SOLVER, [&] {
                launch_extrap_order<R>(stream, n, field_acc.p, grad_acc.p, P.phidp.p, frac_transform<double>(), selfc(), labpol.p,
                                       ptmu.p + m * n3, ptE.p + m * n3, ptG.p + m * n6, ptmu.p + (m + 1) * n3,
                                       P.mu.p);
            });
            check_launch();
            stats.matvecs++;
        }
        timed(P_SOLVER, [&] { launch_extrap_combine<R>(stream, n, ext_cp, ptmu.p, mud.p, P.mu.p); });
        check_launch();
        mu_set = true;
    }

    // DIIS, the reference's solver for mutual polarization
    // (convergeInduceDipolesByDIIS, MPIDReferenceForce.cpp:1182-1291), on the
    // same operator, for parity of the iterates with the reference rather
    // than for speed (MPID_FLAG_DIIS; one host synchronisation per
    // iteration).  In the variables x (mu = S x) the fixed point
    // mu = alpha (E0 + T mu) reads x = b + (I - A) x: the new dipoles are
    // x + r and their error is S r (r = b - A x), the reference's nd - mu,
    // whose RMS in Debye is the stopping test.  The next x combines the last
    // (at most 20) new dipoles with the coefficients of the least-squares
    // combination of their errors (diis_coefficients).  Stops at iteration
    // it with the current x when the error is below the target or it reaches
    // the iteration limit, as the reference; r is then that x's residual.
    static constexpr int kDiisHistory = 20;
    DBuf<double> diis_d, diis_e, diis_dots;
    template <class R> double diis(Prec<R>& P, int& iters) {
        const int n3 = 3 * n;
        const int nslot = kDiisHistory + 1;  // a new entry is added before the oldest leaves
        diis_d.resize((size_t)nslot * n3);
        diis_e.resize((size_t)nslot * n3);
        diis_dots.resize(nslot);
        std::vector<int> hist;  // slots, oldest first
        std::vector<double> B((size_t)nslot * nslot, 0.0), dots(nslot);
        for (int it = 0;; it++) {
            matvec(P, x.p, Ap.p);
            int slot = 0;
            while (std::find(hist.begin(), hist.end(), slot) != hist.end()) slot++;
            hist.push_back(slot);
            double* d = diis_d.p + (size_t)slot * n3;
            double* e = diis_e.p + (size_t)slot * n3;
            timed(P_SOLVER, [&] {
                launch_sub(stream, n3, b.p, Ap.p, r.p);
                HIP_CHECK(hipMemcpyAsync(d, x.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream));
                launch_axpy(stream, n3, 1.0, r.p, d);
                launch_apply_S<double>(stream, n, sqrtpol.p, r.p, e, nullptr);
                for (int s2 : hist)
                    launch_dot(stream, n3, e, diis_e.p + (size_t)s2 * n3, partial.p, diis_dots.p + s2);
            });
            check_launch();
            HIP_CHECK(hipMemcpyAsync(dots.data(), diis_dots.p, nslot * sizeof(double), hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            check_lists_pending();
            const double eps = debye_eps(dots[slot]);
            if (eps < opt.target_epsilon || it == opt.max_iterations) {
                iters = it;
                return eps;
            }
            for (int s2 : hist) B[(size_t)slot * nslot + s2] = B[(size_t)s2 * nslot + slot] = dots[s2];
            if ((int)hist.size() > kDiisHistory) hist.erase(hist.begin());
            const std::vector<double> c = diis_coefficients(B, nslot, hist);
            timed(P_SOLVER, [&] {
                launch_fill<double>(stream, x.p, n3, 0.0);
                for (size_t j = 0; j < hist.size(); j++)
                    launch_axpy(stream, n3, c[j], diis_d.p + (size_t)hist[j] * n3, x.p);
            });
            check_launch();
        }
    }
    // Coefficients c (sum 1) of the entries hist of the error dot products B
    // (nslot x nslot, by slot) that minimise |sum c_j e_j|: the first column,
    // negated, of the pseudo-inverse of the bordered matrix
    // [[0, -1...], [-1, B]], as the reference takes it from JAMA's SVD
    // (MPIDReferenceForce.cpp:1254-1291); here from a cyclic Jacobi
    // eigen-decomposition of the symmetric matrix, with the SVD's rank
    // tolerance.
    static std::vector<double> diis_coefficients(const std::vector<double>& B, int nslot, const std::vector<int>& hist) {
        const int m = (int)hist.size();
        std::vector<double> c(m, 0.0);
        if (m == 1) {
            c[0] = 1.0;
            return c;
        }
        const int R = m + 1;
        std::vector<double> a((size_t)R * R, 0.0), V((size_t)R * R, 0.0);
        for (int i = 0; i < m; i++) a[(size_t)(i + 1) * R] = a[i + 1] = -1.0;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) a[(size_t)(i + 1) * R + j + 1] = B[(size_t)hist[i] * nslot + hist[j]];
        for (int i = 0; i < R; i++) V[(size_t)i * R + i] = 1.0;
        for (int sweep = 0; sweep < 100; sweep++) {
            double off = 0.0;
            for (int p = 0; p < R; p++)
                for (int q = p + 1; q < R; q++) off += a[(size_t)p * R + q] * a[(size_t)p * R + q];
            if (off < 1e-300) break;
            for (int p = 0; p < R; p++)
                for (int q = p + 1; q < R; q++) {
                    const double apq = a[(size_t)p * R + q];
                    if (std::fabs(apq) < 1e-300) continue;
                    const double th = (a[(size_t)q * R + q] - a[(size_t)p * R + p]) / (2.0 * apq);
                    const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                    const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                    for (int k = 0; k < R; k++) {
                        const double akp = a[(size_t)k * R + p], akq = a[(size_t)k * R + q];
                        a[(size_t)k * R + p] = cs * akp - sn * akq;
                        a[(size_t)k * R + q] = sn * akp + cs * akq;
                    }
                    for (int k = 0; k < R; k++) {
                        const double apk = a[(size_t)p * R + k], aqk = a[(size_t)q * R + k];
                        a[(size_t)p * R + k] = cs * apk - sn * aqk;
                        a[(size_t)q * R + k] = sn * apk + cs * aqk;
                    }
                    for (int k = 0; k < R; k++) {
                        const double vkp = V[(size_t)k * R + p], vkq = V[(size_t)k * R + q];
                        V[(size_t)k * R + p] = cs * vkp - sn * vkq;
                        V[(size_t)k * R + q] = sn * vkp + cs * vkq;
                    }
                }
        }
        double smax = 0.0;
        for (int i = 0; i < R; i++) smax = std::max(smax, std::fabs(a[(size_t)i * R + i]));
        const double tol = R * smax * 2.220446049250313e-16;
        for (int i = 1; i < R; i++) {
            double d = 0.0;
            for (int j = 0; j < R; j++) {
                const double w = a[(size_t)j * R + j];
                if (std::fabs(w) > tol) d -= V[j] * V[(size_t)i * R + j] / w;
            }
            c[i - 1] = d;
        }
        return c;
    }

    // residual with the FP64 operator: rt = b - A x; returns its RMS-Debye norm
    double true_residual() {
        const int n3 = 3 * n;
        prof_override = P_REFINE;
        matvec(pd, x.p, Ap.p);
        prof_override = -1;
        timed(P_SOLVER, [&] {
            launch_sub(stream, n3, b.p, Ap.p, rt.p);
            launch_residual_norm(stream, n, sqrtpol.p, rt.p, partial.p, scal.p + 4);
        });
        return debye_eps(read_scalar(scal.p + 5));
    }

    template <class R> void solve(Prec<R>& P) {
        stats.iterations = 0;
        stats.refinements = 0;
        stats.matvecs = 0;
        timed(P_SOLVER, [&] { launch_solver_init(stream, n, sqrtpol.p, E0.p, b.p, x.p); });
        check_launch();
        energy_correction = 0.0;
        correction_pending = false;
        if (opt.polarization == MPID_POLARIZATION_DIRECT || opt.polarization == MPID_POLARIZATION_EXTRAPOLATED) {
            if (opt.polarization == MPID_POLARIZATION_EXTRAPOLATED) extrapolate(P);
            if (lists_pending) {
                wait_stream();
                check_lists_pending();
            }
            stats.converged = 1;
            stats.epsilon = 0.0;
            return;
        }
        int iters = 0;
        double eps = (opt.flags & MPID_FLAG_DIIS) ? diis(P, iters)
                                                 : cg(P, x.p, b.p, false, opt.target_epsilon, iters, opt.max_iterations);
        // Variational correction of the energy: the energy expression of the
        // reference is linear in the induced dipoles, U(mu) = E(mu) - eps/2 y.r
        // (r = b - A y) is stationary at the solution, so it carries no
        // first-order error from the solver tolerance.  It is read with the
        // energy at the end of the evaluation (forces_and_energy).
        auto correction_dot = [&](const double* res) {
            timed(P_SOLVER, [&] { launch_dot(stream, 3 * n, x.p, res, partial.p, scal.p + 6); });
            HIP_CHECK(hipMemcpyAsync(hscal + 2, scal.p + 6, sizeof(double), hipMemcpyDeviceToHost, stream));
        };
        if (opt.precision == MPID_PRECISION_MIXED) {
            // FP64 iterative refinement: the residual of the FP32 solution is
            // formed with the FP64 operator, the correction is solved with the
            // FP32 operator and added in FP64, until the target is met.  The
            // final pass needs only the solution: it is issued first, and the
            // FP64 residual runs on vstream beside it.  Outputs are written
            // only once the check has passed (forces_and_energy); a
            // correction issues the final pass again.
            mark_v();
            if (!defer_final) issue_final_pass(P);
            on_vstream([&] { eps = true_residual(); });
            int refine = 0;
            while (eps >= opt.target_epsilon && refine < 8 && iters < opt.max_iterations) {
                join();    // the final pass's reciprocal chain has read mu
                join_v();  // rt
                refine++;
                const int before = iters;
                cg(P, dx.p, rt.p, true, std::max(0.3 * opt.target_epsilon, 1e-4 * eps), iters, opt.max_iterations);
                if (iters == before) iters++;
                timed(P_SOLVER, [&] { launch_axpy(stream, 3 * n, 1.0, dx.p, x.p); });
                mark_v();
                if (!defer_final) issue_final_pass(P);
                on_vstream([&] { eps = true_residual(); });
            }
            stats.refinements = refine;
            on_vstream([&] { correction_dot(rt.p); });
            join_v();
        } else {
            correction_dot(r.p);
        }
        correction_pending = true;
        stats.iterations = iters;
        stats.epsilon = eps;
        stats.converged = eps < opt.target_epsilon ? 1 : 0;
        if (!stats.converged) {
            char buf[160];
            std::snprintf(buf, sizeof(buf), "Induced dipoles did not converge:  iterations=%d eps=%g", iters, eps);
            throw MpidError(MPID_ERR_NOT_CONVERGED, buf);
        }
    }

    // The final pass: the reciprocal chain of the induced dipoles, the
    // reciprocal per-atom terms and the covalent pairs run on rstream beside
    // the real-space electrostatics; all of them add into the fixed-point
    // force, torque and energy accumulators (integer adds, so the order does
    // not matter).  energy_acc[0, kEnergySlots) takes the real-space energy, the next kEnergySlots the
    // reciprocal and self terms.
    template <class R> void issue_final_pass(Prec<R>& P) {
        join();  // a previous final pass's reciprocal kernels have added into the accumulators
        // induced dipoles: mu = S x (double for the per-atom terms, R for the
        // kernels), unless they are in place already (Extrapolated)
        if (!mu_set) timed(P_SOLVER, [&] { launch_apply_S<R>(stream, n, sqrtpol.p, x.p, P.mu.p, mud.p); });
        timed(P_OTHER, [&] {
            launch_zero_acc(stream, force_acc.p, force_acc.n);
            launch_zero_acc(stream, torque_acc.p, torque_acc.n);
            launch_zero_acc(stream, energy_acc.p, 2 * kEnergySlots);
        });
        if (opt.polarization == MPID_POLARIZATION_EXTRAPOLATED)
            timed(P_SOLVER, [&] {
                launch_extrap_response(stream, n, ext_cp, ptmu.p, ptE.p, ptG.p, aniso.p, force_acc.p, torque_acc.p);
            });
        // Order of issue: the reciprocal chain first (short kernels with large
        // LDS blocks, which would wait for whole CUs behind the
        // electrostatics), then the electrostatics, then the per-atom
        // reciprocal terms and covalent pairs behind the chain on rstream.
        recip_cycle_induced(P, true);  // phidp
        // this category's events bracket the one kernel (its live time in the benchmark)
        timed(P_ELEC, [&] {
            launch_electrostatics_pairs<R>(stream, tile_bound(), tile_act.p, brow_ustart.p + (n + kTile - 1) / kTile,
                                           brows.p, brow_owner.p,
                                           tmask.p, tmaskcc.p, chg.p,
                                           pair_atoms(P), P.mu.p, pair_params<R>(), pd.pos.p, pd.lab.p,
                                           pair_params<double>(), force_acc.p, torque_acc.p, energy_acc.p);
        });
        timed_on(rstream, P_RECIP, [&] {
            launch_recip_atoms<R>(rstream, n, opt.polarization == MPID_POLARIZATION_MUTUAL ? 1 : 0, pd.phi.p,
                                  P.phidp.p, lab.p, mud.p, aniso.p, frac_transform<double>(), opt.ewald_alpha, force_acc.p,
                                  torque_acc.p, energy_acc.p + kEnergySlots);
        });
        timed_on(rstream, P_COV, [&] {
            launch_cov_electrostatics<double>(rstream, n, pair_atoms(pd), mud.p, pair_params<double>(), force_acc.p,
                                              torque_acc.p, energy_acc.p);
        });
        check_launch();
        final_issued = true;
    }

    template <class R> double forces_and_energy(Prec<R>& P, double* d_forces_out) {
        if (!final_issued) issue_final_pass(P);
        final_issued = false;
        mu_valid = true;
        join();
        timed(P_OTHER, [&] {
            launch_torque_to_force(stream, n, dbox, spos.p, saxes.p, torque_acc.p, force_acc.p,
                                   (opt.flags & MPID_FLAG_CONSERVATIVE_TORQUES) ? 1 : 0);
            if (d_forces_out) launch_output_forces(stream, n, perm.p, force_acc.p, d_forces_out, 1);
        });
        check_launch();
        acc_t e[2] = {0, 0};
        acc_t slots[2 * kEnergySlots];
        HIP_CHECK(hipMemcpyAsync(slots, energy_acc.p, sizeof(slots), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        for (int k = 0; k < kEnergySlots; k++) {  // integer sums: any order gives the same bits
            e[0] += slots[k];
            e[1] += slots[kEnergySlots + k];
        }
        if (correction_pending) {
            energy_correction = -0.5 * kElectric * hscal[2];
            correction_pending = false;
        }
        stats.energy_terms[0] = acc_get(e[0]);
        stats.energy_terms[1] = acc_get(e[1]);
        stats.energy_terms[2] = energy_correction;
        return acc_get(e[0]) + acc_get(e[1]) + energy_correction;
    }

    // Lists, permanent field and induced dipoles at d_pos (a list overflow
    // grows the lists and starts over).
    template <class R> void converge(Prec<R>& P, const double* d_pos) {
        for (int attempt = 0;; attempt++) {
            try {
                final_issued = false;
                mu_set = false;
                prepare(d_pos);
                fixed_field(P);
                solve(P);
                break;
            } catch (const ListsOverflow&) {
                if (attempt > 0) throw MpidError(MPID_ERR_HIP, "neighbour list overflow after regrowth");
                // a solver iteration may still run on the old lists: let it
                // finish before the lists are reallocated
                HIP_CHECK(hipStreamSynchronize(stream));
                HIP_CHECK(hipStreamSynchronize(rstream));
                grow_rows(rows_grow_to);
            }
        }
    }
    template <class R> double run(Prec<R>& P, const double* d_pos, double* d_forces_out) {
        converge(P, d_pos);
        double e = forces_and_energy(P, d_forces_out);
        collect_profile();
        return e;
    }

    // ---- energy gradients -------------------------------------------------------
    // The energy of mutual polarization is the variational form
    // U(mu) = E_perm - eps mu.E0 + eps/2 mu.(alpha^-1 - T) mu (what the
    // evaluation returns, stationary in mu), so at the solution its total
    // derivative in any parameter is the partial derivative at fixed mu.  The
    // adjoint solve A lambda = dU/dmu = -r (A = I - S T S, the symmetric
    // operator of the solver, r its residual at the forward solution) moves mu
    // onto the solution to second order in the solver's tolerance; the
    // gradients are then formed at mu + lambda:
    //  positions   minus the forces of the final pass, with the conservative
    //              torques (the exact derivative for anisotropic sites);
    //  moments     the potential vectors dU/dm (pair kernels, reciprocal and
    //              self terms), taken back to the molecular frame;
    //  alpha       -1/2 eps E E^T projected on the molecular axes (E the total
    //              field), plus the damping factors through the Thole terms;
    //  Thole       the derivatives of the pair terms in their damping.
    // Direct polarization (mu = alpha E0, no solve) is U with A = alpha^-1.
    DBuf<acc_t> pot_acc, hd_acc, hdef_acc, ht_acc, gdef_acc;
    DBuf<double> Etot, gmol, gpol, gthole, gforce;
    int adjoint_iterations = 0;
    void alloc_gradients() {
        const size_t n3 = 3 * (size_t)n;
        pot_acc.resize((size_t)kMP * n);
        hd_acc.resize(n);
        hdef_acc.resize(n);
        ht_acc.resize(n);
        gdef_acc.resize(1);
        Etot.resize(n3);
        gmol.resize((size_t)kMP * n);
        gpol.resize(n3);
        gthole.resize(n);
        gforce.resize(n3);
    }
    struct GradientsOut {
        std::vector<double> force, mol, pol, thole;
        double default_thole = 0.0;
    };
    template <class R> double gradients(Prec<R>& P, const double* d_pos, bool all_sites, GradientsOut& out) {
        if (opt.polarization == MPID_POLARIZATION_EXTRAPOLATED)
            throw MpidError(MPID_ERR_UNSUPPORTED, "energy gradients: Extrapolated polarization is not supported "
                                                  "(its energy is not stationary in the induced dipoles)");
        alloc_gradients();
        const int n3 = 3 * n;
        const bool mutual = opt.polarization == MPID_POLARIZATION_MUTUAL;
        // conservative torques for this evaluation; the flag and the deferred
        // final pass are restored however the evaluation ends
        struct Restore {
            mpid_context* c;
            int flags;
            ~Restore() {
                c->opt.flags = flags;
                c->defer_final = false;
            }
        } restore{this, opt.flags};
        opt.flags |= MPID_FLAG_CONSERVATIVE_TORQUES;
        defer_final = true;
        converge(P, d_pos);
        defer_final = false;
        correction_pending = false;  // the energy is taken at the refined dipoles below
        adjoint_iterations = 0;
        if (mutual) {
            // right-hand side: the FP64 residual of the solution (mixed: formed
            // by the check of the solution; FP64: the solver's residual)
            if (opt.precision != MPID_PRECISION_MIXED)
                HIP_CHECK(hipMemcpyAsync(rt.p, r.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream));
            const double stop = 0.01 * stats.epsilon;
            if (stop > 0.0) {
                int it = 0;
                cg(P, dx.p, rt.p, true, stop, it, 20);
                adjoint_iterations = it;
                timed(P_SOLVER, [&] { launch_axpy(stream, n3, 1.0, dx.p, x.p); });
            }
        }
        // final pass at the refined dipoles, with the potential vectors
        timed(P_SOLVER, [&] { launch_apply_S<R>(stream, n, sqrtpol.p, x.p, P.mu.p, mud.p); });
        timed(P_OTHER, [&] {
            launch_zero_acc(stream, force_acc.p, force_acc.n);
            launch_zero_acc(stream, torque_acc.p, torque_acc.n);
            launch_zero_acc(stream, energy_acc.p, 2 * kEnergySlots);
            launch_zero_acc(stream, pot_acc.p, pot_acc.n);
            launch_zero_acc(stream, hd_acc.p, hd_acc.n);
            launch_zero_acc(stream, hdef_acc.p, hdef_acc.n);
            launch_zero_acc(stream, ht_acc.p, ht_acc.n);
            launch_zero_acc(stream, gdef_acc.p, gdef_acc.n);
        });
        recip_cycle_induced(P, true);  // phidp at every atom
        timed(P_ELEC, [&] {
            launch_electrostatics_pot<R>(stream, tile_bound(), tile_act.p, brow_ustart.p + (n + kTile - 1) / kTile,
                                           brows.p, brow_owner.p,
                                         tmask.p, tmaskcc.p, chg.p,
                                         pair_atoms(P), P.mu.p, pair_params<R>(), pd.pos.p, pd.lab.p,
                                         pair_params<double>(), force_acc.p, torque_acc.p, energy_acc.p, pot_acc.p,
                                         all_sites ? 1 : 0);
        });
        timed_on(rstream, P_RECIP, [&] {
            launch_recip_atoms<R>(rstream, n, mutual ? 1 : 0, pd.phi.p, P.phidp.p, lab.p, mud.p, aniso.p, frac_transform<double>(),
                                  opt.ewald_alpha, force_acc.p, torque_acc.p, energy_acc.p + kEnergySlots);
            launch_recip_potential<R>(rstream, n, pme() ? 1 : 0, pd.phi.p, P.phidp.p, lab.p, mud.p, frac_transform<double>(),
                                      opt.ewald_alpha, pot_acc.p);
        });
        timed_on(rstream, P_COV, [&] {
            launch_cov_electrostatics<double>(rstream, n, pair_atoms(pd), mud.p, pair_params<double>(), force_acc.p,
                                              torque_acc.p, energy_acc.p, pot_acc.p);
        });
        // Thole damping of the pairs of polarisable sites
        timed(P_INDUCED, [&] {
            launch_nl_thole<R>(stream, npol, pol_idx.p, nlist_pol(), pair_atoms(P), P.mu.p, pair_params<R>(),
                               hd_acc.p, hdef_acc.p);
        });
        timed(P_COV, [&] {
            launch_cov_thole<double>(stream, n, pair_atoms(pd), mud.p, pair_params<double>(), hd_acc.p, hdef_acc.p,
                                     ht_acc.p);
        });
        // the total field: at the polarisable sites from their lists, or at
        // every site from lists of all atoms (all_sites), E0 included
        if (all_sites) {
            join();  // the reciprocal kernels above read the lab-frame state that prepare() rewrites
            prepare(d_pos, true, false);
            fixed_field(P, true);
        }
        if (mutual) {
            timed(P_INDUCED, [&] {
                launch_zero_acc(stream, field_acc.p, field_acc.n);
                if (all_sites)
                    launch_nl_induced_field<R>(stream, n, nullptr, nlist(), pair_atoms(P), P.mu.p, pair_params<R>(),
                                               field_acc.p);
                else
                    launch_nl_induced_field<R>(stream, npol, pol_idx.p, nlist_pol(), pair_atoms(P), P.mu.p,
                                               pair_params<R>(), field_acc.p);
            });
            timed(P_COV, [&] {
                launch_cov_induced_field<R>(stream, n, pair_atoms(P), P.mu.p, pair_params<R>(), all_sites ? 0 : 1,
                                            field_acc.p);
            });
        }
        join();
        timed(P_OTHER, [&] {
            launch_total_field<R>(stream, n, mutual ? 1 : 0, E0.p, field_acc.p, P.phidp.p, frac_transform<double>(), mud.p, selfc(),
                                  Etot.p);
            launch_torque_to_force(stream, n, dbox, spos.p, saxes.p, torque_acc.p, force_acc.p,
                                   (opt.flags & MPID_FLAG_CONSERVATIVE_TORQUES) ? 1 : 0);
            launch_output_forces(stream, n, perm.p, force_acc.p, gforce.p, 0);
            launch_param_gradients(stream, n, all_sites ? 1 : 0, perm.p, dbox, spos.p, saxes.p, smol.p, spol.p,
                                   damp.p, chg.p, pot_acc.p, Etot.p, hd_acc.p, hdef_acc.p, ht_acc.p,
                                   opt.default_thole, gmol.p, gpol.p, gthole.p, gdef_acc.p);
        });
        check_launch();
        out.force.resize(n3);
        out.mol.resize((size_t)kMP * n);
        out.pol.resize(n3);
        out.thole.resize(n);
        acc_t slots[2 * kEnergySlots];
        acc_t gdef = 0;
        HIP_CHECK(hipMemcpyAsync(slots, energy_acc.p, sizeof(slots), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(&gdef, gdef_acc.p, sizeof(acc_t), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(out.force.data(), gforce.p, n3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(out.mol.data(), gmol.p, out.mol.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 stream));
        HIP_CHECK(hipMemcpyAsync(out.pol.data(), gpol.p, n3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(out.thole.data(), gthole.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        acc_t e[2] = {0, 0};
        for (int k = 0; k < kEnergySlots; k++) {
            e[0] += slots[k];
            e[1] += slots[kEnergySlots + k];
        }
        out.default_thole = acc_get(gdef);
        stats.energy_terms[0] = acc_get(e[0]);
        stats.energy_terms[1] = acc_get(e[1]);
        stats.energy_terms[2] = 0.0;
        mu_valid = true;
        collect_profile();
        return acc_get(e[0]) + acc_get(e[1]);
    }

    // Every path that rewrites the device state forgets the last evaluation
    // first: a getter must not serve a cached result that a failed
    // evaluation has partly overwritten (positions, permutation, dipoles).
    void invalidate_last() {
        have_last = false;
        mu_valid = false;
    }

    double execute_device(const double* d_pos, const double* boxv, double* d_forces) {
        invalidate_last();
        set_box(boxv);
        double e = opt.precision == MPID_PRECISION_MIXED ? run(pf, d_pos, d_forces) : run(pd, d_pos, d_forces);
        return e;
    }

    void remember(const double* pos) {
        last_pos.assign(pos, pos + 3 * (size_t)n);
        last_box.assign(box_vec, box_vec + 9);
        have_last = true;
    }
    bool same_as_last(const double* pos, const double* boxv) const {
        if (!have_last || !mu_valid) return false;
        double b9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // a null box (NoCutoff) reads as zeros
        if (boxv) std::memcpy(b9, boxv, sizeof(b9));
        return std::memcmp(last_pos.data(), pos, 3 * (size_t)n * sizeof(double)) == 0 &&
               std::memcmp(last_box.data(), b9, sizeof(b9)) == 0;
    }

    double execute_host(const double* pos, const double* boxv, double* forces_inout) {
        set_box(boxv);
        pos_in.upload(pos, 3 * (size_t)n, stream);
        if (forces_inout) {
            fout.upload(forces_inout, 3 * (size_t)n, stream);
        }
        double e = execute_device(pos_in.p, boxv, forces_inout ? fout.p : nullptr);
        if (forces_inout) {
            HIP_CHECK(hipMemcpyAsync(forces_inout, fout.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                                     stream));
            HIP_CHECK(hipStreamSynchronize(stream));
        }
        remember(pos);
        return e;
    }

    void ensure(const double* pos, const double* boxv) {
        if (!same_as_last(pos, boxv)) execute_host(pos, boxv, nullptr);
    }
    void download_unpermuted(const double* d_sorted, double* out) {
        launch_unpermute3(stream, n, perm.p, d_sorted, tmp3n.p);
        check_launch();
        HIP_CHECK(hipMemcpyAsync(out, tmp3n.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }
    void lab_dipoles(double* out) {
        std::vector<double> l((size_t)n * kMP), d(3 * (size_t)n);
        std::vector<int> pm(n);
        HIP_CHECK(hipMemcpyAsync(l.data(), lab.p, l.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(pm.data(), perm.p, n * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        for (int s = 0; s < n; s++)
            for (int k = 0; k < 3; k++) out[3 * (size_t)pm[s] + k] = l[(size_t)s * kMP + 1 + k];
    }

    // getElectrostaticPotential after an evaluation at pos (ensure())
    void potential_at_points(int npts, const double* points, double* out) {
        DBuf<double> dpts, dout;
        dpts.upload(points, 3 * (size_t)npts, stream);
        dout.resize(std::max(1, npts));
        launch_potential_points(stream, n, npts, spos.p, lab.p, mud.p, dbox, dpts.p, dout.p);
        check_launch();
        HIP_CHECK(hipMemcpyAsync(out, dout.p, (size_t)npts * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }
    // getSystemMultipoleMoments after an evaluation at the positions in pos_in
    // (MPIDReferenceForce.cpp:2415-2461 for the final traceless form and units)
    void multipole_moments(const double* masses, double* out) {
        DBuf<double> dm, dsum;
        dm.upload(masses, n, stream);
        dsum.resize(20);
        launch_multipole_moments(stream, n, perm.p, pos_in.p, dm.p, lab.p, mud.p, dsum.p);
        check_launch();
        double v[20];
        HIP_CHECK(hipMemcpyAsync(v, dsum.p, sizeof(v), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        const double* t = v + 8;   // xx xy xz yy yz zz
        const double* qa = v + 14; // atomic quadrupoles, same order
        const double tr = (t[0] + t[3] + t[5]) / 3.0;
        const double debye = 4.80321;  // the reference's value, :2447
        out[0] = v[4];
        for (int k = 0; k < 3; k++) out[1 + k] = v[5 + k] * 10.0 * debye;
        const int at[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
                const int k = at[a][c];
                out[4 + 3 * a + c] = (0.5 * (t[k] - (a == c ? tr : 0.0)) + qa[k]) * 100.0 * 3.0 * debye;
            }
    }

    static_assert(kNoAllocMaxPoints == MPID_NOALLOC_MAX_POINTS, "staging limit of the header");
    // getElectrostaticPotential of the last completed evaluation without any
    // allocation: points and values pass through the per-device staging array
    // (noalloc_scratch), under one lock, since every context on a device shares
    // that array.
    void potential_at_points_noalloc(int npts, const double* points, double* out) {
        if (!mu_valid)
            throw MpidError(MPID_ERR_INVALID, "getElectrostaticPotential (allocation-free): the context holds no "
                                              "completed evaluation");
        if (npts > kNoAllocMaxPoints)
            throw MpidError(MPID_ERR_INVALID, "getElectrostaticPotential (allocation-free): " + std::to_string(npts) +
                                                  " points exceed the limit of " +
                                                  std::to_string(kNoAllocMaxPoints) + " per call");
        if (!noalloc_scratch) throw MpidError(MPID_ERR_HIP, "the staging array's device address is unknown");
        if (npts == 0) return;
        static std::mutex lock;
        std::lock_guard<std::mutex> g(lock);
        double* dpts = noalloc_scratch;
        double* dval = noalloc_scratch + 3 * (size_t)kNoAllocMaxPoints;
        HIP_CHECK(hipMemcpyAsync(dpts, points, 3 * (size_t)npts * sizeof(double), hipMemcpyHostToDevice, stream));
        launch_potential_points(stream, n, npts, spos.p, lab.p, mud.p, dbox, dpts, dval);
        check_launch();
        HIP_CHECK(hipMemcpyAsync(out, dval, (size_t)npts * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    double energy_gradients(const double* pos, const double* boxv, bool all_sites, GradientsOut& out) {
        invalidate_last();
        set_box(boxv);
        pos_in.upload(pos, 3 * (size_t)n, stream);
        const double e = opt.precision == MPID_PRECISION_MIXED ? gradients(pf, pos_in.p, all_sites, out)
                                                               : gradients(pd, pos_in.p, all_sites, out);
        remember(pos);
        return e;
    }

    void compute_fixed_field(const double* pos, const double* boxv, double* out) {
        invalidate_last();
        set_box(boxv);
        pos_in.upload(pos, 3 * (size_t)n, stream);
        prepare(pos_in.p, true);
        if (opt.precision == MPID_PRECISION_MIXED) fixed_field(pf, true); else fixed_field(pd, true);
        download_unpermuted(E0.p, out);
        have_last = false;
        collect_profile();
    }
    void compute_induced_field(const double* pos, const double* boxv, const double* mu, double* out) {
        invalidate_last();
        set_box(boxv);
        pos_in.upload(pos, 3 * (size_t)n, stream);
        prepare(pos_in.p, true, false);
        // mu (original order) -> sorted precision copy
        std::vector<int> pm(n);
        HIP_CHECK(hipMemcpyAsync(pm.data(), perm.p, n * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        auto run_field = [&](auto& P) {
            using R = typename std::remove_reference<decltype(*P.mu.p)>::type;
            std::vector<R> ms(3 * (size_t)n);
            for (int s = 0; s < n; s++)
                for (int k = 0; k < 3; k++) ms[3 * (size_t)s + k] = (R)mu[3 * (size_t)pm[s] + k];
            HIP_CHECK(hipMemcpyAsync(P.mu.p, ms.data(), ms.size() * sizeof(R), hipMemcpyHostToDevice, stream));
            induced_field(P, true);
            join();
            launch_field_sum<R>(stream, n, field_acc.p, P.erec.p, P.mu.p, selfc(), tmp3n.p);
            check_launch();
            std::vector<double> f(3 * (size_t)n);
            HIP_CHECK(hipMemcpyAsync(f.data(), tmp3n.p, f.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            for (int s = 0; s < n; s++)
                for (int k = 0; k < 3; k++) out[3 * (size_t)pm[s] + k] = f[3 * (size_t)s + k];
        };
        if (opt.precision == MPID_PRECISION_MIXED) run_field(pf); else run_field(pd);
        have_last = false;
        collect_profile();
    }
};

namespace {
thread_local std::string g_create_error;

template <class F> int guarded(mpid_context* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->error.clear();
        return MPID_OK;
    } catch (const MpidError& e) {
        if (ctx) ctx->error = e.what();
        return e.code;
    } catch (const HipError& e) {
        if (ctx) ctx->error = e.what();
        return MPID_ERR_HIP;
    } catch (const std::exception& e) {
        if (ctx) ctx->error = e.what();
        return MPID_ERR_HIP;
    }
}
}  // namespace

extern "C" {

int mpid_create(const mpid_parameters* params, const mpid_options* opts, mpid_context** out) {
    if (!out) return MPID_ERR_INVALID;
    *out = nullptr;
    std::unique_ptr<mpid_context> ctx(new mpid_context());
    int rc = guarded(ctx.get(), [&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw MpidError(MPID_ERR_HIP, "no HIP device is available");
        HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        if (opts && (opts->flags & MPID_FLAG_SINGLE_STREAM)) {
            ctx->rstream = ctx->vstream = ctx->stream;
        } else {
            HIP_CHECK(hipStreamCreateWithFlags(&ctx->rstream, hipStreamNonBlocking));
            HIP_CHECK(hipStreamCreateWithFlags(&ctx->vstream, hipStreamNonBlocking));
        }
        HIP_CHECK(hipEventCreateWithFlags(&ctx->ev_v, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ctx->ev_vdone, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ctx->ev_sync, hipEventDisableTiming));
        HIP_CHECK(hipHostMalloc((void**)&ctx->hscal, 16 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
        for (hipEvent_t& e : ctx->ev_check) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIP_CHECK(hipHostGetDevicePointer((void**)&ctx->hscal_dev, ctx->hscal, 0));

        HIP_CHECK(hipHostMalloc((void**)&ctx->hnl, 4 * sizeof(int), 0));
        ctx->set_options(opts);
        ctx->set_parameters(params, true);
        ctx->alloc_state();
        // the allocation-free getter's staging array: its address on this
        // device, looked up once per device with the device current
        HIP_CHECK(noalloc_scratch_address(&ctx->noalloc_scratch));
    });
    if (rc != MPID_OK) {
        g_create_error = ctx->error;
        return rc;
    }
    *out = ctx.release();
    return MPID_OK;
}

const char* mpid_create_error(void) { return g_create_error.c_str(); }

int mpid_update_parameters(mpid_context* ctx, const mpid_parameters* params) {
    if (!ctx) return MPID_ERR_INVALID;
    return guarded(ctx, [&] { ctx->set_parameters(params, false); });
}

int mpid_execute(mpid_context* ctx, const double* positions, const double* box, int include_forces,
                 int include_energy, double* forces_inout, double* energy_out) {
    if (!ctx || !positions) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        double e = ctx->execute_host(positions, box, include_forces ? forces_inout : nullptr);
        if (energy_out && include_energy) *energy_out = e;
        else if (energy_out) *energy_out = 0.0;
    });
}

int mpid_execute_device(mpid_context* ctx, const double* d_positions, const double* box, double* d_forces_inout,
                        double* energy_out) {
    if (!ctx || !d_positions) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        double e = ctx->execute_device(d_positions, box, d_forces_inout);
        ctx->have_last = false;
        if (energy_out) *energy_out = e;
    });
}

int mpid_get_electrostatic_potential(mpid_context* ctx, const double* positions, const double* box, int npoints,
                                     const double* points, double* out) {
    if (!ctx || !positions || npoints < 0 || (npoints > 0 && (!points || !out))) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        ctx->ensure(positions, box);
        if (npoints > 0) ctx->potential_at_points(npoints, points, out);
    });
}

int mpid_get_electrostatic_potential_noalloc(mpid_context* ctx, int npoints, const double* points, double* out) {
    if (!ctx || npoints < 0 || (npoints > 0 && (!points || !out))) return MPID_ERR_INVALID;
    return guarded(ctx, [&] { ctx->potential_at_points_noalloc(npoints, points, out); });
}

int mpid_energy_gradients(mpid_context* ctx, const double* positions, const double* box, int flags,
                          double* energy_out, mpid_gradients* grads) {
    if (!ctx || !positions || !grads) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        mpid_context::GradientsOut g;
        const double e = ctx->energy_gradients(positions, box, (flags & MPID_GRADIENT_ALL_SITES) != 0, g);
        const int n = ctx->n;
        if (energy_out) *energy_out = e;
        grads->adjoint_iterations = ctx->adjoint_iterations;
        for (int i = 0; i < n; i++) {
            const double* m = &g.mol[(size_t)kMP * i];
            if (grads->positions)
                for (int k = 0; k < 3; k++) grads->positions[3 * (size_t)i + k] = -g.force[3 * (size_t)i + k];
            if (grads->charges) grads->charges[i] = m[0];
            if (grads->dipoles)
                for (int k = 0; k < 3; k++) grads->dipoles[3 * (size_t)i + k] = m[1 + k];
            // internal XX XY XZ YY YZ ZZ -> MPIDForce XX XY YY XZ YZ ZZ (the
            // inverse of the ordering map in set_parameters)
            if (grads->quadrupoles) {
                double* q = grads->quadrupoles + 6 * (size_t)i;
                q[0] = m[4]; q[1] = m[5]; q[3] = m[6]; q[2] = m[7]; q[4] = m[8]; q[5] = m[9];
            }
            if (grads->octopoles) {
                double* o = grads->octopoles + 10 * (size_t)i;
                o[0] = m[10]; o[1] = m[11]; o[4] = m[12]; o[2] = m[13]; o[5] = m[14];
                o[7] = m[15]; o[3] = m[16]; o[6] = m[17]; o[8] = m[18]; o[9] = m[19];
            }
            if (grads->polarity)
                for (int k = 0; k < 3; k++) grads->polarity[3 * (size_t)i + k] = g.pol[3 * (size_t)i + k];
            if (grads->tholes) grads->tholes[i] = g.thole[i];
        }
        if (grads->default_thole) *grads->default_thole = g.default_thole;
    });
}

int mpid_get_system_multipole_moments(mpid_context* ctx, const double* positions, const double* box,
                                      const double* masses, double* out) {
    if (!ctx || !positions || !masses || !out) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        ctx->ensure(positions, box);
        ctx->multipole_moments(masses, out);
    });
}

int mpid_get_induced_dipoles(mpid_context* ctx, const double* positions, const double* box, double* out) {
    if (!ctx || !positions || !out) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        ctx->ensure(positions, box);
        ctx->download_unpermuted(ctx->mud.p, out);
    });
}

int mpid_get_lab_frame_permanent_dipoles(mpid_context* ctx, const double* positions, const double* box, double* out) {
    if (!ctx || !positions || !out) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        ctx->ensure(positions, box);
        ctx->lab_dipoles(out);
    });
}

int mpid_get_total_dipoles(mpid_context* ctx, const double* positions, const double* box, double* out) {
    if (!ctx || !positions || !out) return MPID_ERR_INVALID;
    return guarded(ctx, [&] {
        ctx->ensure(positions, box);
        std::vector<double> ind(3 * (size_t)ctx->n);
        ctx->download_unpermuted(ctx->mud.p, ind.data());
        ctx->lab_dipoles(out);
        for (size_t k = 0; k < ind.size(); k++) out[k] += ind[k];
    });
}

int mpid_get_pme_bspline_order(void) { return kOrder; }

int mpid_spread_tilings(const int* grid, int max, int* levels, int* modes) {
    if (!grid || max < 0 || (max > 0 && (!levels || !modes))) return -MPID_ERR_INVALID;
    for (int k = 0; k < 3; k++) {
        int rr[16], c;
        if (grid[k] < kOrder || grid[k] > 1024 || !factor_radices(grid[k], rr, c)) return -MPID_ERR_INVALID;
    }
    const std::vector<SpreadTiling> t = spread_tilings(grid);
    for (int i = 0; i < (int)t.size() && i < max; i++) {
        levels[i] = t[i].level;
        modes[i] = t[i].mode;
    }
    return (int)t.size();
}

int mpid_get_pme_parameters(const mpid_context* ctx, double* alpha, int* nx, int* ny, int* nz) {
    if (!ctx) return MPID_ERR_INVALID;
    if (alpha) *alpha = ctx->opt.ewald_alpha;
    if (nx) *nx = ctx->opt.grid[0];
    if (ny) *ny = ctx->opt.grid[1];
    if (nz) *nz = ctx->opt.grid[2];
    return MPID_OK;
}

int mpid_compute_fixed_field(mpid_context* ctx, const double* positions, const double* box, double* field_out) {
    if (!ctx || !positions || !field_out) return MPID_ERR_INVALID;
    return guarded(ctx, [&] { ctx->compute_fixed_field(positions, box, field_out); });
}

int mpid_compute_induced_field(mpid_context* ctx, const double* positions, const double* box, const double* mu,
                               double* field_out) {
    if (!ctx || !positions || !mu || !field_out) return MPID_ERR_INVALID;
    return guarded(ctx, [&] { ctx->compute_induced_field(positions, box, mu, field_out); });
}

int mpid_get_stats(const mpid_context* ctx, mpid_stats* stats) {
    if (!ctx || !stats) return MPID_ERR_INVALID;
    *stats = ctx->stats;
    return MPID_OK;
}

int mpid_set_profile_mask(mpid_context* ctx, unsigned mask) {
    if (!ctx) return MPID_ERR_INVALID;
    ctx->prof_mask = mask & ((1u << MPID_PROF_COUNT) - 1u);
    return MPID_OK;
}
int mpid_set_profiling(mpid_context* ctx, int enable) {
    return mpid_set_profile_mask(ctx, enable ? ~0u : 0u);
}

int mpid_get_profile(mpid_context* ctx, double* times_ms, int* counts, int reset) {
    if (!ctx) return MPID_ERR_INVALID;
    for (int k = 0; k < MPID_PROF_COUNT; k++) {
        if (times_ms) times_ms[k] = ctx->prof_ms[k];
        if (counts) counts[k] = ctx->prof_count[k];
        if (reset) {
            ctx->prof_ms[k] = 0;
            ctx->prof_count[k] = 0;
        }
    }
    return MPID_OK;
}

const char* mpid_last_error(const mpid_context* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

void mpid_destroy(mpid_context* ctx) { delete ctx; }

}  // extern "C"
