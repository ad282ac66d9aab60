/*
 * C ABI of the MI355X (gfx950) MPIDForce PME + induced-dipole evaluator.
 *
 * This is the drop-in boundary that replaces the body of the reference's
 * platform kernel, class CalcMPIDForceKernel
 * (openmmapi/include/openmm/mpidKernels.h:50-104).  Each entry point names the
 * virtual method it stands in for.  Plain pointers and sizes only; all device
 * memory is owned by the context.  Not thread safe per context; distinct
 * contexts may be used from distinct threads.
 *
 * Orderings are those of the MPIDForce API (openmmapi/include/openmm/MPIDForce.h:236-237):
 *   dipole      X Y Z
 *   quadrupole  XX XY YY XZ YZ ZZ
 *   octopole    XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ
 * Units: nm, e, kJ/mol.  Positions are [n][3] doubles; box is three box
 * vectors a, b, c, row-major [3][3], in OpenMM's reduced form (a along x, b
 * in the xy plane, |bx|, |cx| <= ax/2, |cy| <= by/2; rectangular or
 * triclinic, MPIDReferenceForce.cpp:2615-2676): other boxes are
 * MPID_ERR_INVALID.
 */
#ifndef MPID_HIP_H
#define MPID_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpid_context mpid_context;

/* status codes (0 = success) */
enum {
    MPID_OK = 0,
    MPID_ERR_INVALID = 1,        /* invalid parameter (OpenMMException in the reference) */
    MPID_ERR_NOT_CONVERGED = 2,  /* mutual induced dipoles did not converge */
    MPID_ERR_HIP = 3,            /* HIP runtime failure */
    MPID_ERR_UNSUPPORTED = 4     /* feature outside this build (NoCutoff beyond 16,384 particles, ...) */
};

/* NOCUTOFF: every pair, no periodic images, no reciprocal space; the cutoff,
 * Ewald and grid settings and the box are ignored; at most 16,384 particles. */
enum { MPID_NONBONDED_NOCUTOFF = 0, MPID_NONBONDED_PME = 1 };
enum { MPID_POLARIZATION_MUTUAL = 0, MPID_POLARIZATION_DIRECT = 1, MPID_POLARIZATION_EXTRAPOLATED = 2 };
/* MIXED: FP32 pair and reciprocal arithmetic, 64-bit fixed-point accumulation,
 * FP64 solver reductions and FP64 iterative refinement of the induced dipoles.
 * DOUBLE: the same kernels in FP64 throughout. */
enum { MPID_PRECISION_MIXED = 0, MPID_PRECISION_DOUBLE = 1 };
/* Option flags.
 * CONSERVATIVE_TORQUES: add the half of the induced-induced torque about the
 * pair axis that the reference omits (MPIDReferenceForce.cpp:4854-4855), and
 * map the torques of ThreeFold sites onto their frame atoms by the exact
 * chain rule of the frame instead of the reference's approximate projection
 * (:2068-2097).  Off by default, which reproduces the reference; with it the
 * forces are exact gradients of the energy for every site.
 * SINGLE_STREAM: issue all of the context's work on one HIP stream instead of
 * three (real space, reciprocal space, the mixed path's FP64 check).  For many
 * small contexts driven at once, where cross-stream waits would hold up the
 * hardware queues the contexts share; results are the same.
 * DIIS: solve mutual polarization with the reference's DIIS iteration
 * (convergeInduceDipolesByDIIS, MPIDReferenceForce.cpp:1182-1291: history of
 * 20, coefficients from the pseudo-inverse of the bordered error matrix,
 * the same Debye-RMS stopping test) instead of preconditioned conjugate
 * gradients, so that the iterates and the iteration count follow the
 * reference's; one host synchronisation per iteration.  Double precision
 * only (MPID_ERR_UNSUPPORTED otherwise). */
enum { MPID_FLAG_CONSERVATIVE_TORQUES = 1, MPID_FLAG_SINGLE_STREAM = 2, MPID_FLAG_DIIS = 4 };

/* Global settings of the force (MPIDForce.h:118-470 setters). */
typedef struct {
    int nonbonded_method;   /* MPID_NONBONDED_PME or _NOCUTOFF */
    int polarization;       /* MPID_POLARIZATION_MUTUAL, _DIRECT or _EXTRAPOLATED */
    double cutoff;          /* nm */
    double ewald_alpha;     /* nm^-1, must be > 0 */
    int grid[3];            /* PME grid, sizes of the form 2^a 3^b 5^c */
    double default_thole;   /* MPIDForce::getDefaultTholeWidth */
    double scale14;         /* MPIDForce::get14ScaleFactor */
    int max_iterations;     /* MPIDForce::getMutualInducedMaxIterations */
    double target_epsilon;  /* MPIDForce::getMutualInducedTargetEpsilon (RMS Debye) */
    int precision;          /* MPID_PRECISION_* */
    int flags;              /* MPID_FLAG_* bits, 0 for the reference behaviour */
    /* Extrapolated polarization: MPIDForce::setExtrapolationCoefficients
     * (MPIDForce.h:346; copied to the reference force at
     * MPIDReferenceKernels.cpp:149-150, 213-215).  n_extrapolation = 0 takes
     * MPIDForce's defaults (-0.154, 0.017, 0.658, 0.474; MPIDForce.cpp:46-49);
     * at most MPID_MAX_EXTRAPOLATION coefficients.  The array is copied. */
    int n_extrapolation;
    const double* extrapolation_coefficients;
    /* Tiling of the PME spreading (no counterpart in the reference, whose
     * spreading is one loop over the atoms): one of the pairs that
     * mpid_spread_tilings lists for the grid, as an autotuner picks them, or
     * spread_level = 0 for the built-in choice.  spread_level -1: no tiles
     * (atomic adds into the grid); m >= 1: each grid axis cut into 2^m tiles,
     * with spread_mode 1 (each workgroup owns a tile and writes it once) or
     * 2 (overlapping windows summed into the grid afterwards). */
    int spread_level;
    int spread_mode;
} mpid_options;
enum { MPID_MAX_EXTRAPOLATION = 16 };

/* Per-particle parameters (MPIDForce::addMultipole, setCovalentMap).
 * Covalent lists in CSR form: the atoms in list t (0 = Covalent12, 1 = 13,
 * 2 = 14, 3 = 15) of atom a are
 * covalent_atoms[covalent_offsets[4a+t] .. covalent_offsets[4a+t+1]). */
typedef struct {
    int n;
    const double* charges;       /* [n]   */
    const double* dipoles;       /* [3n]  molecular frame */
    const double* quadrupoles;   /* [6n]  molecular frame */
    const double* octopoles;     /* [10n] molecular frame */
    const int* axis_types;       /* [n]   MPIDForce::MultipoleAxisTypes */
    const int* atom_z;           /* [n]   */
    const int* atom_x;           /* [n]   */
    const int* atom_y;           /* [n]   */
    const double* tholes;        /* [n]   */
    const double* polarity;      /* [3n]  molecular-frame xx yy zz */
    const int* covalent_offsets; /* [4n+1] */
    const int* covalent_atoms;
} mpid_parameters;

typedef struct {
    int iterations;          /* solver iterations of the last execute */
    double epsilon;          /* final RMS induced-dipole residual, Debye */
    int converged;
    int refinements;         /* FP64 refinement passes (mixed precision) */
    int matvecs;             /* matrix-vector products */
    int ntiles;              /* 64x64 block pairs within the cutoff */
    double energy_terms[3];  /* real space; reciprocal + self; variational solver correction */
} mpid_stats;

/* Per-kernel-family device time of the executes since profiling was enabled
 * (HIP events on the stream that runs the family's kernels; the reciprocal
 * chains run on a second stream beside the real-space kernels, so the
 * families' times can overlap).  MPID_PROF_FFT covers the whole reciprocal
 * convolution (the influence function is applied inside the FFT passes), so
 * MPID_PROF_CONVOLVE stays zero.  MPID_PROF_ELECTROSTATIC_PAIRS times the
 * one real-space electrostatics kernel alone. */
enum {
    MPID_PROF_SORT = 0, MPID_PROF_ROTATE, MPID_PROF_TILES, MPID_PROF_SPREAD, MPID_PROF_FFT, MPID_PROF_CONVOLVE,
    MPID_PROF_GATHER, MPID_PROF_FIXED_PAIRS, MPID_PROF_INDUCED_PAIRS, MPID_PROF_ELECTROSTATIC_PAIRS,
    MPID_PROF_RECIP_ATOMS, MPID_PROF_SOLVER, MPID_PROF_OTHER,
    MPID_PROF_COVALENT_PAIRS, /* the scaled covalent pairs (FP64 permanent field and electrostatics) */
    MPID_PROF_REFINEMENT,     /* mixed precision: the FP64 operator applications of the refinement */
    MPID_PROF_COUNT
};

/* CalcMPIDForceKernel::initialize (mpidKernels.h:67).  Validates as
 * MPIDForceImpl::initialize (openmmapi/src/MPIDForceImpl.cpp:51-150). */
int mpid_create(const mpid_parameters* params, const mpid_options* opts, mpid_context** out);

/* CalcMPIDForceKernel::copyParametersToContext (mpidKernels.h:93).  The
 * particle count must not change. */
int mpid_update_parameters(mpid_context* ctx, const mpid_parameters* params);

/* CalcMPIDForceKernel::execute (mpidKernels.h:77): host positions; box: the
 * three box vectors, row-major (may be NULL for NoCutoff); forces are ADDED
 * into forces_inout ([n][3], may be NULL); energy (kJ/mol) written to
 * *energy_out (may be NULL). */
int mpid_execute(mpid_context* ctx, const double* positions, const double* box, int include_forces,
                 int include_energy, double* forces_inout, double* energy_out);

/* Same, with positions and forces in device memory (HIP pointers on the
 * context's device).  The call is synchronous. */
int mpid_execute_device(mpid_context* ctx, const double* d_positions, const double* box, double* d_forces_inout,
                        double* energy_out);

/* CalcMPIDForceKernel::getInducedDipoles / getLabFramePermanentDipoles /
 * getTotalDipoles (mpidKernels.h:79-81): [n][3] e nm at the given positions
 * (re-evaluated when they differ from the last execute). */
int mpid_get_induced_dipoles(mpid_context* ctx, const double* positions, const double* box, double* out);
/* CalcMPIDForceKernel::getElectrostaticPotential (mpidKernels.h:83-84): at
 * npoints points ([npoints][3], nm) the potential (kJ/mol/e) of every
 * particle's charge, dipole plus induced dipole and quadrupole (octopoles
 * are not included, as MPIDReferenceForce.cpp:2464-2489), the separation
 * taken to its nearest image under PME, after an evaluation at positions. */
int mpid_get_electrostatic_potential(mpid_context* ctx, const double* positions, const double* box, int npoints,
                                     const double* points, double* out);
/* Allocation-free form of mpid_get_electrostatic_potential.  It has no
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It returns the potential of the context's last completed
 * evaluation (mpid_execute, mpid_execute_device or a getter), with host arrays
 * in and out, and calls no hipMalloc, hipFree or allocating routine.  Points
 * and values are staged in one fixed device array per GPU (2 MiB, held until
 * the process ends), so a call takes at most MPID_NOALLOC_MAX_POINTS points:
 * more fail with MPID_ERR_INVALID, as does a context without a completed
 * evaluation.  Every context on one GPU shares that array; calls within a
 * process are serialised on it, and the array keeps the last call's data. */
enum { MPID_NOALLOC_MAX_POINTS = 65536 };
int mpid_get_electrostatic_potential_noalloc(mpid_context* ctx, int npoints, const double* points, double* out);
/* CalcMPIDForceKernel::getSystemMultipoleMoments (mpidKernels.h:86): 13
 * values about the centre of mass of masses ([n], the System's particle
 * masses, MPIDReferenceKernels.cpp:325-337): net charge (e), dipole (Debye,
 * with the induced dipoles) and the traceless quadrupole (Buckingham, 3x3
 * row-major) as MPIDReferenceForce.cpp:2352-2462, after an evaluation at
 * positions. */
int mpid_get_system_multipole_moments(mpid_context* ctx, const double* positions, const double* box,
                                      const double* masses, double* out13);
int mpid_get_lab_frame_permanent_dipoles(mpid_context* ctx, const double* positions, const double* box, double* out);
int mpid_get_total_dipoles(mpid_context* ctx, const double* positions, const double* box, double* out);

/* The spreading tilings available for a PME grid of grid[0..2] points, as
 * (spread_level, spread_mode) pairs for mpid_options, the built-in choice
 * first.  No counterpart in the reference: it is for an autotuner, which
 * times each.  Writes at most max pairs and returns how many there are, or
 * -MPID_ERR_INVALID (negative) for a grid the PME path does not take. */
int mpid_spread_tilings(const int* grid, int max, int* levels, int* modes);

/* MPIDForce::getPmeBSplineOrder (MPIDForce.h:196): the B-spline order this
 * library was built for, 6 (as the reference) or 5 (libmpid_hip_order5.so,
 * built with -DMPID_PME_ORDER=5). */
int mpid_get_pme_bspline_order(void);

/* CalcMPIDForceKernel::getPMEParameters (mpidKernels.h:103). */
int mpid_get_pme_parameters(const mpid_context* ctx, double* alpha, int* nx, int* ny, int* nz);

/* Inspection of intermediate results for parity tests:
 * permanent field E (before the polarisability), [n][3]; and the field
 * T*mu (real space + reciprocal + self) of given induced dipoles, [n][3]. */
int mpid_compute_fixed_field(mpid_context* ctx, const double* positions, const double* box, double* field_out);
int mpid_compute_induced_field(mpid_context* ctx, const double* positions, const double* box, const double* mu,
                               double* field_out);

/* Gradients of the energy with respect to the positions and the force's
 * per-particle parameters.  No counterpart in the reference, which returns
 * forces only; the interface it extends is CalcMPIDForceKernel::execute
 * (mpidKernels.h:77) and the parameters of MPIDForce::addMultipole
 * (MPIDForce.h:236-281).
 *
 * One evaluation at positions (mutual or direct polarization; Extrapolated
 * returns MPID_ERR_UNSUPPORTED), with the conservative torques whatever the
 * context's flags, followed by an adjoint solve with the solver's symmetric
 * operator, whose right-hand side is the derivative of the energy in the
 * induced dipoles (the residual of the solution).  The gradients are those of
 * the energy at the refined dipoles, which *energy_out receives.  Each output
 * pointer may be NULL.  Gradients are with respect to the numbers in the
 * layouts of mpid_parameters, each listed component separately (a
 * quadrupole's XY once, not XY and YX).
 *
 * Sites with neither higher moments nor polarizability ("bare" charges) get
 * the charge gradient only, and zeros for their other parameters, unless
 * flags has MPID_GRADIENT_ALL_SITES, which evaluates every pair with the
 * general formula and the total field at every site (slower). */
typedef struct {
    double* positions;      /* [n][3] dE/dr (kJ/mol/nm): minus the forces */
    double* charges;        /* [n] kJ/mol/e */
    double* dipoles;        /* [3n] molecular frame */
    double* quadrupoles;    /* [6n] molecular frame, XX XY YY XZ YZ ZZ */
    double* octopoles;      /* [10n] molecular frame, XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ */
    double* polarity;       /* [3n] molecular-frame xx yy zz polarizabilities (kJ/mol/nm^3) */
    double* tholes;         /* [n] */
    double* default_thole;  /* [1] MPIDForce::getDefaultTholeWidth */
    int adjoint_iterations; /* out: iterations of the adjoint solve */
} mpid_gradients;
enum { MPID_GRADIENT_ALL_SITES = 1 };
int mpid_energy_gradients(mpid_context* ctx, const double* positions, const double* box, int flags,
                          double* energy_out, mpid_gradients* grads);

int mpid_get_stats(const mpid_context* ctx, mpid_stats* stats);
/* Profiling (no counterpart in the reference): mpid_set_profiling times every
 * family (enable != 0) or none; mpid_set_profile_mask times the families
 * whose bit (1 << MPID_PROF_x) is set.  When a family issued inside a solver
 * iteration is timed (spread, FFT, gather, induced pairs, covalent pairs,
 * solver), the iterations run as individual launches; otherwise they are
 * replayed from a captured graph, as without profiling. */
int mpid_set_profiling(mpid_context* ctx, int enable);
int mpid_set_profile_mask(mpid_context* ctx, unsigned mask);
/* times_ms[MPID_PROF_COUNT], counts[MPID_PROF_COUNT] (launches); resets them when reset != 0 */
int mpid_get_profile(mpid_context* ctx, double* times_ms, int* counts, int reset);

const char* mpid_last_error(const mpid_context* ctx);
/* error of the last failed mpid_create on this thread */
const char* mpid_create_error(void);
void mpid_destroy(mpid_context* ctx);

#ifdef __cplusplus
}
#endif
#endif
