/*
 * FP64 CPU oracle for the MPIDForce PME + induced-dipole path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product path (mpidopenmmplugin_amd/) never links or calls it.
 *
 * It restates, in plain double precision, the algorithm of
 * platforms/reference/src/SimTKReference/MPIDReferenceForce.cpp
 * (class MPIDReferencePmeForce) together with the parameter flattening of
 * platforms/reference/src/MPIDReferenceKernels.cpp:84-239 and the validation
 * of openmmapi/src/MPIDForceImpl.cpp:51-150.  Each function in
 * mpid_oracle.cpp cites the reference lines it follows.
 */
#ifndef MPID_ORACLE_H
#define MPID_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Input orderings are the MPIDForce API orderings (MPIDForce.h:236-237):
 * quadrupole XX XY YY XZ YZ ZZ, octopole XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ.
 * Covalent lists are CSR: the atoms of list type t (0=1-2, 1=1-3, 2=1-4, 3=1-5)
 * of atom a are cov_atoms[cov_offsets[4*a+t] .. cov_offsets[4*a+t+1]). */
typedef struct {
    int n;
    const double* charges;      /* [n]    */
    const double* dipoles;      /* [3n]   */
    const double* quadrupoles;  /* [6n]   */
    const double* octopoles;    /* [10n]  */
    const int* axis_types;      /* [n]    */
    const int* atom_z;          /* [n]    */
    const int* atom_x;          /* [n]    */
    const int* atom_y;          /* [n]    */
    const double* tholes;       /* [n]    */
    const double* polarity;     /* [3n]   */
    const int* cov_offsets;     /* [4n+1] */
    const int* cov_atoms;
    int polarization;           /* 0 = Mutual, 1 = Direct, 2 = Extrapolated */
    double cutoff;
    double alpha;
    int nx, ny, nz;
    double default_thole;
    double scale14;
    int max_iter;
    double target_eps;
    double box[9];              /* three box vectors, row-major */
    int flags;                  /* 1: restore the omitted induced-induced z torque */
    int n_coeff;                /* Extrapolated: number of coefficients (>= 1) */
    const double* coeffs;       /* Extrapolated: MPIDForce::getExtrapolationCoefficients */
    int method;                 /* 0 = NoCutoff (all pairs, no periodicity, no Ewald), 1 = PME */
} oracle_system;

/* Every pointer may be NULL.  Internal orderings: quadrupole QXX QXY QXZ QYY
 * QYZ QZZ, octopole QXXX QXXY QXXZ QXYY QXYZ QXZZ QYYY QYYZ QYZZ QZZZ,
 * polarisation tensor XX XY XZ YY YZ ZZ. */
typedef struct {
    double* lab_dipoles;     /* [3n]  */
    double* lab_quads;       /* [6n]  */
    double* lab_octs;        /* [10n] */
    double* lab_pol;         /* [6n]  */
    double* fixed_field;     /* [3n]  permanent field E (before alpha) */
    double* induced;         /* [3n]  converged induced dipoles */
    double* phi;             /* [35n] fractional potential derivatives, fixed grid */
    double* phidp;           /* [35n] fractional potential derivatives, induced grid */
    double* torques;         /* [3n]  */
    double* forces;          /* [3n]  overwritten (not accumulated) */
    double energy_terms[4];  /* real space, recip fixed, recip induced, self */
    double energy;
    /* Mutual polarization: the energy plus -1/2 eps mu.(E0 + T mu - alpha^-1 mu),
     * which vanishes at the exact solution and makes the value stationary in
     * the dipoles (its error is second order in the solver's); otherwise the
     * energy itself.  For finite differences in the parameters. */
    double energy_variational;
    int iterations;
    double eps;
    int converged;
    char error[256];
} oracle_out;

/* Solver modes: 0 = DIIS as the reference (MPIDReferenceForce.cpp:1182-1291). */
int oracle_evaluate(const oracle_system* sys, const double* pos, oracle_out* out);

/* Building blocks, exposed for unit tests. */
/* theta[order][5]: the weights of the order (6, or 5 in the order-5 build)
 * grid points and their first four derivatives */
void oracle_bspline(double w, double* theta);
void oracle_bspline_moduli(int n, double* moduli);
int oracle_fft3d(double* data_interleaved, int nx, int ny, int nz, int sign);
/* Induced field T*mu (real + reciprocal + self) for given lab dipoles, no
 * solver: the matrix-vector product of the mutual solver. */
int oracle_induced_field(const oracle_system* sys, const double* pos, const double* mu, double* field_out);
/* mapTorqueToForce (MPIDReferenceForce.cpp:1895-2131) alone: forces [3n]
 * (overwritten) from per-site torques [3n] at positions [3n]; only n, the
 * axis types and the frame atoms of sys are read. */
int oracle_torque_to_force(const oracle_system* sys, const double* pos, const double* torques, double* forces);
/* getElectrostaticPotential after the solve (MPIDReferenceForce.cpp:2464-2531):
 * at each of the npts points, the Coulomb constant times the bare potential
 * of every particle's charge, dipole plus induced dipole and quadrupole
 * (octopoles are not included), nearest image under PME. */
int oracle_potential(const oracle_system* sys, const double* pos, int npts, const double* points, double* out);
/* getSystemMultipoleMoments after the solve (MPIDReferenceForce.cpp:2352-2462):
 * net charge, dipole (Debye) and traceless quadrupole (Buckingham), 13
 * values, about the centre of mass of the given masses. */
int oracle_multipole_moments(const oracle_system* sys, const double* pos, const double* masses, double* out13);

#ifdef __cplusplus
}
#endif
#endif
