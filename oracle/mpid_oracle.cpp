// FP64 CPU oracle for the MPIDForce PME path (Mutual, Direct and Extrapolated polarization).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker.  The product library never
// links or calls this file.
//
// This is a restatement of MPIDReferencePmeForce
// (platforms/reference/src/SimTKReference/MPIDReferenceForce.cpp) in plain
// double precision.  Where the reference hand-expands a tensor identity
// (fractional transforms, Cartesian potential transform, reciprocal torques,
// spherical rotation matrices) this file evaluates the same identity in its
// generic tensor form, so that a transcription slip in one of the two cannot
// hide behind the other; the golden values of the reference's own tests pin
// both.  Deliberate deviations, each noted at its site:
//   * NoAxisType sites are not rotated (the reference rotates them as ZThenX
//     whenever multipoleAtomZ >= 0; the CUDA platform uses identity).
//   * ZOnly torque mapping picks the second frame vector as the rotation does
//     (the reference indexes atom -1 there; torques on the reference tests'
//     ZOnly sites are zero, so the goldens are unaffected).
//   * DIIS coefficients use a symmetric Jacobi eigen-decomposition of the
//     B matrix, which is the same pseudo-inverse the reference takes from
//     JAMA's SVD.
#include "mpid_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#include <cstdlib>

// One line on stderr per solver iteration when MPID_ORACLE_PROGRESS is set.
static bool progress_lines() {
    static const bool on = std::getenv("MPID_ORACLE_PROGRESS") != nullptr;
    return on;
}

namespace {

typedef std::complex<double> cplx;

// MPIDReferenceForce.cpp:39 (Coulomb constant of the reference force) and :49 (Debye).
const double kElectric = 138.935455846;
const double kDebye = 48.033324;
// MPIDReferenceForce.cpp:2545 keeps a truncated sqrt(pi); kept for parity.
const double kSqrtPi = 1.77245385091;
// MPIDReferenceForce.cpp:2543, B-spline order hard-wired to 6; 5 in a build
// with -DMPID_PME_ORDER=5 (libmpid_oracle_order5.so), the same formulas
#ifndef MPID_PME_ORDER
#define MPID_PME_ORDER 6
#endif
const int kOrder = MPID_PME_ORDER;
static_assert(kOrder == 5 || kOrder == 6, "B-spline order 5 or 6");

struct V3 {
    double x[3];
    V3() { x[0] = x[1] = x[2] = 0.0; }
    V3(double a, double b, double c) { x[0] = a; x[1] = b; x[2] = c; }
    double& operator[](int i) { return x[i]; }
    double operator[](int i) const { return x[i]; }
    V3 operator+(const V3& o) const { return V3(x[0] + o[0], x[1] + o[1], x[2] + o[2]); }
    V3 operator-(const V3& o) const { return V3(x[0] - o[0], x[1] - o[1], x[2] - o[2]); }
    V3 operator*(double s) const { return V3(x[0] * s, x[1] * s, x[2] * s); }
    V3& operator+=(const V3& o) { x[0] += o[0]; x[1] += o[1]; x[2] += o[2]; return *this; }
    V3& operator-=(const V3& o) { x[0] -= o[0]; x[1] -= o[1]; x[2] -= o[2]; return *this; }
    double dot(const V3& o) const { return x[0] * o[0] + x[1] * o[1] + x[2] * o[2]; }
    V3 cross(const V3& o) const {
        return V3(x[1] * o[2] - x[2] * o[1], x[2] * o[0] - x[0] * o[2], x[0] * o[1] - x[1] * o[0]);
    }
};

double normalize(V3& v) {
    double n = std::sqrt(v.dot(v));
    if (n > 0.0) v = v * (1.0 / n);
    return n;
}

// ---------------------------------------------------------------------------
// Full symmetric Cartesian tensors of rank 2 and 3 in the reference's internal
// orderings (MPIDReferenceForce.h:809-810).
// ---------------------------------------------------------------------------
int idx2(int a, int b) {
    static const int t[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return t[a][b];
}
int idx3(int a, int b, int c) {
    int s[3] = {a, b, c};
    std::sort(s, s + 3);
    static const int t[3][3][3] = {
        {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}},
        {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
        {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
    return t[s[0]][s[1]][s[2]];
}
// Levi-Civita symbol.
int eps3(int a, int b, int c) {
    if (a == b || b == c || a == c) return 0;
    return ((a == 0 && b == 1) || (a == 1 && b == 2) || (a == 2 && b == 0)) ? 1 : -1;
}

// Spherical-harmonic forms of Cartesian moments, MPIDReferenceForce.cpp:321-342.
void sph_dip(const V3& d, double s[3]) { s[0] = d[2]; s[1] = d[0]; s[2] = d[1]; }
void sph_quad(const double* Q, double s[5]) {
    const double c43 = 2.0 / std::sqrt(3.0), c13 = 1.0 / std::sqrt(3.0);
    s[0] = 3.0 * Q[5];
    s[1] = 3.0 * c43 * Q[2];
    s[2] = 3.0 * c43 * Q[4];
    s[3] = 3.0 * c13 * (Q[0] - Q[3]);
    s[4] = 3.0 * c43 * Q[1];
}
void sph_oct(const double* O, double s[7]) {
    const double r32 = std::sqrt(1.5), r35 = std::sqrt(0.6), r110 = std::sqrt(0.1);
    s[0] = 15.0 * O[9];
    s[1] = 15.0 * r32 * O[5];
    s[2] = 15.0 * r32 * O[8];
    s[3] = 15.0 * r35 * (O[2] - O[7]);
    s[4] = 15.0 * 2.0 * r35 * O[4];
    s[5] = 15.0 * r110 * (O[0] - 3.0 * O[3]);
    s[6] = 15.0 * r110 * (3.0 * O[1] - O[6]);
}

// Rotate tensors into a frame whose axes (in lab components) are e[0..2]:
// T'_{ab..} = sum e[a][i] e[b][j] .. T_{ij..}.
V3 frame_vec(const V3 e[3], const V3& v) { return V3(e[0].dot(v), e[1].dot(v), e[2].dot(v)); }
void frame_quad(const V3 e[3], const double* Q, double* out) {
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            double s = 0;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) s += e[a][i] * e[b][j] * Q[idx2(i, j)];
            out[idx2(a, b)] = s;
        }
}
void frame_oct(const V3 e[3], const double* O, double* out) {
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++)
            for (int c = b; c < 3; c++) {
                double s = 0;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        for (int k = 0; k < 3; k++) s += e[a][i] * e[b][j] * e[c][k] * O[idx3(i, j, k)];
                out[idx3(a, b, c)] = s;
            }
}

// Infinitesimal active rotation about axis k (delta r = e_k x r) applied to
// each tensor slot: the derivative of the moments with respect to the angle.
V3 rot_vec(int k, const V3& d) {
    V3 r;
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < 3; i++) r[a] += eps3(a, k, i) * d[i];
    return r;
}
void rot_quad(int k, const double* Q, double* out) {
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            double s = 0;
            for (int i = 0; i < 3; i++) s += eps3(a, k, i) * Q[idx2(i, b)] + eps3(b, k, i) * Q[idx2(a, i)];
            out[idx2(a, b)] = s;
        }
}
void rot_oct(int k, const double* O, double* out) {
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++)
            for (int c = b; c < 3; c++) {
                double s = 0;
                for (int i = 0; i < 3; i++)
                    s += eps3(a, k, i) * O[idx3(i, b, c)] + eps3(b, k, i) * O[idx3(a, i, c)] +
                         eps3(c, k, i) * O[idx3(a, b, i)];
                out[idx3(a, b, c)] = s;
            }
}
// Full contraction sum over every ordered index tuple of a symmetric tensor
// (stored by unique components) with a second one.
double contract2(const double* A, const double* B) {
    double s = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) s += A[idx2(a, b)] * B[idx2(a, b)];
    return s;
}
double contract3(const double* A, const double* B) {
    double s = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) s += A[idx3(a, b, c)] * B[idx3(a, b, c)];
    return s;
}

struct Site {
    V3 pos;
    double q;
    V3 d;
    double Q[6];
    double O[10];
    double sd[3], sq[5], so[7];
    double thole, damp;
    double pol[3];
    double labpol[6];
    bool aniso;
};

// ---------------------------------------------------------------------------
// B-splines.  Restates computeBSplinePoint, MPIDReferenceForce.cpp:2956-3044:
// order-p cardinal B-spline weights of the 6 grid points by the
// Cox-de Boor recursion, and their r-th derivatives as r-fold backward
// differences of the order-(p-r) weights.
// ---------------------------------------------------------------------------
void bspline(double w, double theta[kOrder][5]) {
    double c[kOrder + 1][kOrder + 1];
    c[1][0] = 1.0;
    for (int p = 2; p <= kOrder; p++) {
        for (int m = 0; m < p; m++) {
            double left = (m >= 1) ? c[p - 1][m - 1] : 0.0;
            double right = (m <= p - 2) ? c[p - 1][m] : 0.0;
            c[p][m] = ((w + (p - 1 - m)) * left + (m + 1 - w) * right) / (p - 1);
        }
    }
    for (int r = 0; r <= 4; r++) {
        int p = kOrder - r;
        double cur[kOrder + 1];
        for (int m = 0; m < p; m++) cur[m] = c[p][m];
        int len = p;
        for (int s = 0; s < r; s++) {
            double nxt[kOrder + 1];
            for (int m = 0; m <= len; m++) nxt[m] = (m >= 1 ? cur[m - 1] : 0.0) - (m < len ? cur[m] : 0.0);
            len++;
            for (int m = 0; m < len; m++) cur[m] = nxt[m];
        }
        for (int m = 0; m < kOrder; m++) theta[m][r] = cur[m];
    }
}

// ---------------------------------------------------------------------------
// FFT: unnormalised complex DFT by recursive mixed-radix decimation in time
// (any length; prime factors above 5 fall back to a direct DFT).  The
// reference calls OpenMM's fftpack (MPIDReferenceForce.cpp:2931-2933); the
// convolution kernel is real and even in k, so the sign convention is
// immaterial.
// ---------------------------------------------------------------------------
// W[k] = exp(sign 2 pi i k / ntop); a sub-transform of length n uses W[k * ntop / n].
void fft1d(cplx* x, int n, int stride, const std::vector<cplx>& W, int ntop) {
    if (n == 1) return;
    int p = 0;
    for (int f : {4, 2, 3, 5})
        if (n % f == 0) { p = f; break; }
    if (p == 0) {
        for (int f = 7; f * f <= n; f += 2)
            if (n % f == 0) { p = f; break; }
        if (p == 0) p = n;
    }
    int m = n / p;
    int step = ntop / n;
    std::vector<cplx> sub(n);
    for (int r = 0; r < p; r++)
        for (int k = 0; k < m; k++) sub[r * m + k] = x[(k * p + r) * stride];
    for (int r = 0; r < p; r++) fft1d(&sub[r * m], m, 1, W, ntop);
    for (int k = 0; k < n; k++) {
        cplx s = 0;
        int kk = k % m;
        for (int r = 0; r < p; r++) s += sub[r * m + kk] * W[(size_t)((long long)r * k % n) * step];
        x[k * stride] = s;
    }
}

std::vector<cplx> twiddles(int n, int sign) {
    std::vector<cplx> W(n);
    for (int k = 0; k < n; k++) {
        double ang = sign * 2.0 * M_PI * k / n;
        W[k] = cplx(std::cos(ang), std::sin(ang));
    }
    return W;
}

void fft3d(std::vector<cplx>& g, int nx, int ny, int nz, int sign) {
    std::vector<cplx> line;
    std::vector<cplx> Wz = twiddles(nz, sign), Wy = twiddles(ny, sign), Wx = twiddles(nx, sign);
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) fft1d(&g[((size_t)x * ny + y) * nz], nz, 1, Wz, nz);
    line.resize(ny);
    for (int x = 0; x < nx; x++)
        for (int z = 0; z < nz; z++) {
            for (int y = 0; y < ny; y++) line[y] = g[((size_t)x * ny + y) * nz + z];
            fft1d(line.data(), ny, 1, Wy, ny);
            for (int y = 0; y < ny; y++) g[((size_t)x * ny + y) * nz + z] = line[y];
        }
    line.resize(nx);
    for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
            for (int x = 0; x < nx; x++) line[x] = g[((size_t)x * ny + y) * nz + z];
            fft1d(line.data(), nx, 1, Wx, nx);
            for (int x = 0; x < nx; x++) g[((size_t)x * ny + y) * nz + z] = line[x];
        }
}

// initializeBSplineModuli, MPIDReferenceForce.cpp:2720-2810.
void bspline_moduli(int n, double* mod) {
    double th[kOrder][5];
    bspline(0.0, th);
    // b(m) for m = 0..n-1: b(0) = 0, b(m) = theta_{m-1}(0) for m = 1..order.
    std::vector<double> b(std::max(n, kOrder + 1), 0.0);
    for (int m = 1; m <= kOrder; m++) b[m] = th[m - 1][0];
    for (int i = 0; i < n; i++) {
        double s1 = 0, s2 = 0;
        for (int m = 0; m < n; m++) {
            double arg = 2.0 * M_PI * i * m / n;
            s1 += b[m] * std::cos(arg);
            s2 += b[m] * std::sin(arg);
        }
        mod[i] = s1 * s1 + s2 * s2;
    }
    const double eps = 1.0e-7;
    if (mod[0] < eps) mod[0] = 0.5 * mod[1];
    for (int i = 1; i < n - 1; i++)
        if (mod[i] < eps) mod[i] = 0.5 * (mod[i - 1] + mod[i + 1]);
    if (mod[n - 1] < eps) mod[n - 1] = 0.5 * mod[n - 2];
    const int jcut = 50;
    for (int i = 1; i <= n; i++) {
        int k = i - 1;
        if (i > n / 2) k -= n;
        double zeta = 1.0;
        if (k != 0) {
            double s1 = 1.0, s2 = 1.0, f = M_PI * k / n;
            for (int j = 1; j <= jcut; j++) {
                double a = f / (f + M_PI * j);
                s1 += std::pow(a, kOrder);
                s2 += std::pow(a, 2 * kOrder);
            }
            for (int j = 1; j <= jcut; j++) {
                double a = f / (f - M_PI * j);
                s1 += std::pow(a, kOrder);
                s2 += std::pow(a, 2 * kOrder);
            }
            zeta = s2 / s1;
        }
        mod[i - 1] *= zeta * zeta;
    }
}

// Order of the 35 fractional potential derivatives (MPIDReferenceForce.cpp:3498-3532).
const int kPhiExp[35][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1},
    {0, 1, 1}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1}, {1, 0, 2},
    {0, 1, 2}, {1, 1, 1}, {4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {3, 1, 0}, {3, 0, 1}, {1, 3, 0}, {0, 3, 1},
    {1, 0, 3}, {0, 1, 3}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};
int phi_index(int a, int b, int c) {
    for (int i = 0; i < 35; i++)
        if (kPhiExp[i][0] == a && kPhiExp[i][1] == b && kPhiExp[i][2] == c) return i;
    return -1;
}

// Multipole component k (0..19) as derivative exponents: q, 3 dipole, 6 quadrupole
// (internal order), 10 octopole (internal order).
void comp_exp(int k, int e[3]) {
    e[0] = e[1] = e[2] = 0;
    if (k == 0) return;
    if (k < 4) { e[k - 1] = 1; return; }
    static const int q[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    if (k < 10) { e[q[k - 4][0]]++; e[q[k - 4][1]]++; return; }
    static const int o[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                                 {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
    for (int s = 0; s < 3; s++) e[o[k - 10][s]]++;
}

struct Oracle {
    oracle_system S;  // a copy: NoCutoff sets its Ewald alpha to zero
    int n;
    std::vector<Site> site;
    V3 box[3], recip[3];
    double A[3][3];  // A[i][j] = n_j * recip[i][j]
    int gn[3];
    std::vector<double> moduli[3];
    std::vector<double> theta[3];  // [n*6*5]
    std::vector<int> igrid;        // [3n]
    std::vector<cplx> grid;
    std::vector<double> mfrac;     // [20n] fractional multipoles
    std::vector<double> phi, phidp;
    std::vector<V3> E0, mu;
    // Extrapolated polarization: the perturbation-order dipoles mu_(m), the
    // field of mu_(m) (with the self term) and its gradient (xx yy zz xy xz yz,
    // without it) for m = 0 .. N-2, and the partial sums of the coefficients.
    std::vector<std::vector<V3>> pt_mu, pt_E;
    std::vector<std::vector<double>> pt_G;
    std::vector<double> cpart;
    // real-space pair list
    struct Pair { int i, j; double mscale, pscale; };
    std::vector<Pair> pairs;
    int iterations = 0;
    double last_eps = 0;
    bool converged = false;

    // NoCutoff (MPIDReferenceForce, the non-PME class): every pair, no
    // periodic images, no reciprocal or self terms.  Its pair formulas are
    // the PME real-space ones at alpha = 0 (erfc = 1, no Gaussian terms);
    // the reference's NoCutoff goldens pin that identity.
    bool periodic = true;

    explicit Oracle(const oracle_system& s) : S(s), n(s.n), periodic(s.method == 1) {
        if (!periodic) S.alpha = 0.0;  // the non-PME reference force has no Ewald screening
    }

    // ---- validation: MPIDForceImpl.cpp:51-150 and MPIDReferenceKernels.cpp:193-197
    void validate() {
        if (S.method != 0 && S.method != 1) throw std::runtime_error("MPIDForce: unknown nonbonded method");
        for (int k = 0; k < 3 && periodic; k++)
            if (S.cutoff > 0.5 * S.box[4 * k]) throw std::runtime_error("MPIDForce: The cutoff distance cannot be greater than half the periodic box size.");
        for (int i = 0; i < n; i++) {
            const double* q = S.quadrupoles + 6 * i;
            const double* o = S.octopoles + 10 * i;
            if (std::fabs(q[0] + q[2] + q[5]) > 1e-5) throw std::runtime_error("MPIDForce: quadrupole has nonzero trace");
            if (std::fabs(o[0] + o[2] + o[7]) > 1e-5 || std::fabs(o[1] + o[3] + o[8]) > 1e-5 ||
                std::fabs(o[4] + o[6] + o[9]) > 1e-5)
                throw std::runtime_error("MPIDForce: octopole has nonzero trace");
            int t = S.axis_types[i];
            if (t < 0 || t > 5) throw std::runtime_error("MPIDForce: axis type not handled");
            if (t != 5 && (S.atom_z[i] < 0 || S.atom_z[i] >= n)) throw std::runtime_error("MPIDForce: invalid z axis particle");
            if (t != 5 && t != 4 && (S.atom_x[i] < 0 || S.atom_x[i] >= n)) throw std::runtime_error("MPIDForce: invalid x axis particle");
            if ((t == 2 || t == 3) && (S.atom_y[i] < 0 || S.atom_y[i] >= n)) throw std::runtime_error("MPIDForce: invalid y axis particle");
        }
        if (S.polarization < 0 || S.polarization > 2) throw std::runtime_error("MPIDForce: unknown polarization type");
        if (S.polarization == 2 && (S.n_coeff < 1 || !S.coeffs))
            throw std::runtime_error("MPIDForce: extrapolated polarization needs at least one coefficient");
    }

    // ---- setPeriodicBoxSize, MPIDReferenceForce.cpp:2615-2633
    void set_box() {
        for (int k = 0; k < 3; k++) box[k] = V3(S.box[3 * k], S.box[3 * k + 1], S.box[3 * k + 2]);
        if (!periodic) {  // no images, no grid: the reciprocal transforms vanish
            for (int k = 0; k < 3; k++) {
                recip[k] = V3();
                gn[k] = 0;
                for (int j = 0; j < 3; j++) A[k][j] = 0.0;
            }
            return;
        }
        double det = box[0][0] * box[1][1] * box[2][2];
        double sc = 1.0 / det;
        recip[0] = V3(box[1][1] * box[2][2], 0, 0) * sc;
        recip[1] = V3(-box[1][0] * box[2][2], box[0][0] * box[2][2], 0) * sc;
        recip[2] = V3(box[1][0] * box[2][1] - box[1][1] * box[2][0], -box[0][0] * box[2][1], box[0][0] * box[1][1]) * sc;
        gn[0] = S.nx; gn[1] = S.ny; gn[2] = S.nz;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = gn[j] * recip[i][j];
    }
    // getPeriodicDelta, MPIDReferenceForce.cpp:2671-2676
    void min_image(V3& d) const {
        if (!periodic) return;
        d -= box[2] * std::floor(d[2] * recip[2][2] + 0.5);
        d -= box[1] * std::floor(d[1] * recip[1][1] + 0.5);
        d -= box[0] * std::floor(d[0] * recip[0][0] + 0.5);
    }

    // ---- loadParticleData (:282-350), checkChiral (:357-397), applyRotationMatrix (:399-648, 788-802)
    void load_and_rotate(const double* pos) {
        site.resize(n);
        for (int i = 0; i < n; i++) {
            Site& s = site[i];
            s.pos = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            s.q = S.charges[i];
            s.d = V3(S.dipoles[3 * i], S.dipoles[3 * i + 1], S.dipoles[3 * i + 2]);
            const double* q = S.quadrupoles + 6 * i;
            // input XX XY YY XZ YZ ZZ -> internal XX XY XZ YY YZ ZZ
            s.Q[0] = q[0]; s.Q[1] = q[1]; s.Q[2] = q[3]; s.Q[3] = q[2]; s.Q[4] = q[4]; s.Q[5] = q[5];
            const double* o = S.octopoles + 10 * i;
            // input XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ -> internal XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ
            s.O[0] = o[0]; s.O[1] = o[1]; s.O[2] = o[4]; s.O[3] = o[2]; s.O[4] = o[5];
            s.O[5] = o[7]; s.O[6] = o[3]; s.O[7] = o[6]; s.O[8] = o[8]; s.O[9] = o[9];
            s.thole = S.tholes[i];
            for (int k = 0; k < 3; k++) s.pol[k] = S.polarity[3 * i + k];
            // MPIDReferenceKernels.cpp:134, dampingFactor = (mean alpha)^(1/6)
            s.damp = std::pow((s.pol[0] + s.pol[1] + s.pol[2]) / 3.0, 1.0 / 6.0);
            s.aniso = s.pol[0] != s.pol[1] || s.pol[0] != s.pol[2];
        }
        // chirality: flips y-odd dipole and quadrupole components only (:357-380)
        for (int i = 0; i < n; i++) {
            if (S.axis_types[i] != 0 || S.atom_y[i] < 0) continue;
            const Site &Z = site[S.atom_z[i]], &X = site[S.atom_x[i]], &Y = site[S.atom_y[i]];
            V3 ad = site[i].pos - Y.pos, bd = Z.pos - Y.pos, cd = X.pos - Y.pos;
            if (bd.cross(cd).dot(ad) < 0.0) {
                site[i].d[1] = -site[i].d[1];
                site[i].Q[1] = -site[i].Q[1];
                site[i].Q[4] = -site[i].Q[4];
            }
        }
        std::vector<V3> P(n);
        for (int i = 0; i < n; i++) P[i] = site[i].pos;
        for (int i = 0; i < n; i++) {
            Site& s = site[i];
            int t = S.axis_types[i];
            V3 ax[3];  // molecular x, y, z axes in lab components
            if (t == 5 || S.atom_z[i] < 0) {
                ax[0] = V3(1, 0, 0); ax[1] = V3(0, 1, 0); ax[2] = V3(0, 0, 1);
            } else {
                V3 vz = P[S.atom_z[i]] - P[i];
                normalize(vz);
                V3 vx;
                if (t == 4) {
                    vx = std::fabs(vz[0]) < 0.866 ? V3(1, 0, 0) : V3(0, 1, 0);
                } else {
                    vx = P[S.atom_x[i]] - P[i];
                    if (t == 1) {
                        normalize(vx);
                        vz += vx;
                        normalize(vz);
                    } else if (t == 2) {
                        normalize(vx);
                        V3 vy = P[S.atom_y[i]] - P[i];
                        normalize(vy);
                        vx += vy;
                        normalize(vx);
                    } else if (t == 3) {
                        normalize(vx);
                        V3 vy = P[S.atom_y[i]] - P[i];
                        normalize(vy);
                        vz += vx + vy;
                        normalize(vz);
                    }
                }
                vx -= vz * vz.dot(vx);
                normalize(vx);
                ax[0] = vx; ax[1] = vz.cross(vx); ax[2] = vz;
            }
            // lab tensor = sum over molecular axes: T_lab = R^T T_mol R with rows of R = ax.
            V3 lt[3];  // lab axes expressed in the molecular frame: lt[a][k] = ax[k][a]
            for (int a = 0; a < 3; a++)
                for (int k = 0; k < 3; k++) lt[a][k] = ax[k][a];
            V3 d = frame_vec(lt, s.d);
            double Q[6], O[10];
            frame_quad(lt, s.Q, Q);
            frame_oct(lt, s.O, O);
            double pm[6] = {s.pol[0], 0, 0, s.pol[1], 0, s.pol[2]};
            frame_quad(lt, pm, s.labpol);
            s.d = d;
            std::memcpy(s.Q, Q, sizeof(Q));
            std::memcpy(s.O, O, sizeof(O));
            sph_dip(s.d, s.sd);
            sph_quad(s.Q, s.sq);
            sph_oct(s.O, s.so);
        }
    }

    // ---- setupScaleMaps / getMultipoleScaleFactor (:190-241): scale of the
    // pair (i<j) comes from the lists of the lower index; later list types
    // overwrite earlier ones; m and p scales are {0, 0, scale14, 1}.
    void scale_of(int i, int j, double& ms, double& ps) const {
        int lo = std::min(i, j), hi = std::max(i, j);
        const double table[4] = {0.0, 0.0, S.scale14, 1.0};
        ms = ps = 1.0;
        for (int t = 0; t < 4; t++)
            for (int k = S.cov_offsets[4 * lo + t]; k < S.cov_offsets[4 * lo + t + 1]; k++)
                if (S.cov_atoms[k] == hi) ms = ps = table[t];
    }

    void build_pairs() {
        pairs.clear();
        double rc2 = S.cutoff * S.cutoff;
        bool rect = box[1][0] == 0 && box[2][0] == 0 && box[2][1] == 0;
        auto consider = [&](int i, int j) {
            V3 d = site[j].pos - site[i].pos;
            min_image(d);
            if (d.dot(d) > rc2) return;
            Pair p;
            p.i = i; p.j = j;
            scale_of(i, j, p.mscale, p.pscale);
            pairs.push_back(p);
        };
        if (!periodic) {
            for (int i = 0; i < n; i++)
                for (int j = i + 1; j < n; j++) {
                    Pair p;
                    p.i = i; p.j = j;
                    scale_of(i, j, p.mscale, p.pscale);
                    pairs.push_back(p);
                }
            return;
        }
        if (n <= 4096 || !rect) {
            for (int i = 0; i < n; i++)
                for (int j = i + 1; j < n; j++) consider(i, j);
            return;
        }
        // cell list over the rectangular box, cells no smaller than the cutoff
        int nc[3];
        for (int k = 0; k < 3; k++) nc[k] = std::max(1, (int)std::floor(box[k][k] / S.cutoff));
        std::vector<std::vector<int>> cells(nc[0] * nc[1] * nc[2]);
        std::vector<int> cid(n);
        for (int i = 0; i < n; i++) {
            int c[3];
            for (int k = 0; k < 3; k++) {
                double f = site[i].pos[k] / box[k][k];
                f -= std::floor(f);
                c[k] = std::min(nc[k] - 1, (int)(f * nc[k]));
            }
            cid[i] = (c[0] * nc[1] + c[1]) * nc[2] + c[2];
            cells[cid[i]].push_back(i);
        }
        for (int i = 0; i < n; i++) {
            int c0 = cid[i] / (nc[1] * nc[2]), c1 = (cid[i] / nc[2]) % nc[1], c2 = cid[i] % nc[2];
            std::vector<int> seen;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++)
                    for (int c = -1; c <= 1; c++) {
                        int x = (c0 + a + nc[0]) % nc[0], y = (c1 + b + nc[1]) % nc[1], z = (c2 + c + nc[2]) % nc[2];
                        seen.push_back((x * nc[1] + y) * nc[2] + z);
                    }
            std::sort(seen.begin(), seen.end());
            seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
            for (int cc : seen)
                for (int j : cells[cc])
                    if (j > i) consider(i, j);
        }
        std::sort(pairs.begin(), pairs.end(), [](const Pair& a, const Pair& b) { return a.i != b.i ? a.i < b.i : a.j < b.j; });
    }

    // ---- PME reciprocal space --------------------------------------------
    // computeMPIDBsplines, :3049-3075
    void compute_splines() {
        theta[0].assign((size_t)n * kOrder * 5, 0.0);
        theta[1] = theta[0];
        theta[2] = theta[0];
        igrid.assign(3 * n, 0);
        for (int i = 0; i < n; i++) {
            V3 p = site[i].pos;
            min_image(p);
            for (int k = 0; k < 3; k++) {
                double w = p[0] * recip[0][k] + p[1] * recip[1][k] + p[2] * recip[2][k];
                double fr = gn[k] * (w - (int)(w + 0.5) + 0.5);
                int ifr = (int)std::floor(fr);
                w = fr - ifr;
                int ig = ifr - kOrder + 1;
                if (ig < 0) ig += gn[k];
                igrid[3 * i + k] = ig;
                double th[kOrder][5];
                bspline(w, th);
                for (int m = 0; m < kOrder; m++)
                    for (int r = 0; r < 5; r++) theta[k][((size_t)i * kOrder + m) * 5 + r] = th[m][r];
            }
        }
    }
    double th(int k, int i, int m, int r) const { return theta[k][((size_t)i * kOrder + m) * 5 + r]; }

    // transformMultipolesToFractionalCoordinates, :3077-3169: full tensor
    // transform by A^T, times the multiplicity of each unique component.
    void fractional_multipoles() {
        mfrac.assign((size_t)20 * n, 0.0);
        for (int i = 0; i < n; i++) {
            const Site& s = site[i];
            double* m = &mfrac[20 * i];
            m[0] = s.q;
            for (int j = 0; j < 3; j++) {
                double v = 0;
                for (int k = 0; k < 3; k++) v += A[k][j] * s.d[k];
                m[1 + j] = v;
            }
            for (int a = 0; a < 3; a++)
                for (int b = a; b < 3; b++) {
                    double v = 0;
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++) v += A[k][a] * A[l][b] * s.Q[idx2(k, l)];
                    m[4 + idx2(a, b)] = v * (a == b ? 1.0 : 2.0);
                }
            for (int a = 0; a < 3; a++)
                for (int b = a; b < 3; b++)
                    for (int c = b; c < 3; c++) {
                        double v = 0;
                        for (int k = 0; k < 3; k++)
                            for (int l = 0; l < 3; l++)
                                for (int o = 0; o < 3; o++) v += A[k][a] * A[l][b] * A[o][c] * s.O[idx3(k, l, o)];
                        int mult = (a == b && b == c) ? 1 : ((a == b || b == c) ? 3 : 6);
                        m[10 + idx3(a, b, c)] = v * mult;
                    }
        }
    }

    size_t gidx(int x, int y, int z) const { return ((size_t)x * gn[1] + y) * gn[2] + z; }

    // spreadFixedMultipolesOntoGrid (:3269-3327) for 20 components, or
    // spreadInducedDipolesOnGrid (:3532-3573) when ncomp == 4 (charge slot 0 unused).
    void spread(const std::vector<double>& m20, int ncomp) {
        grid.assign((size_t)gn[0] * gn[1] * gn[2], cplx(0, 0));
        int ce[20][3];
        for (int k = 0; k < 20; k++) comp_exp(k, ce[k]);
        for (int i = 0; i < n; i++) {
            const double* m = &m20[(size_t)20 * i];
            for (int ix = 0; ix < kOrder; ix++) {
                int x = (igrid[3 * i] + ix) % gn[0];
                for (int iy = 0; iy < kOrder; iy++) {
                    int y = (igrid[3 * i + 1] + iy) % gn[1];
                    for (int iz = 0; iz < kOrder; iz++) {
                        int z = (igrid[3 * i + 2] + iz) % gn[2];
                        double v = 0;
                        for (int k = 0; k < ncomp; k++)
                            v += m[k] * th(0, i, ix, ce[k][0]) * th(1, i, iy, ce[k][1]) * th(2, i, iz, ce[k][2]);
                        grid[gidx(x, y, z)] += v;
                    }
                }
            }
        }
    }

    // performMPIDReciprocalConvolution, :3329-3366
    void convolve() {
        double expf = (M_PI * M_PI) / (S.alpha * S.alpha);
        double sc = 1.0 / (M_PI * box[0][0] * box[1][1] * box[2][2]);
        for (int kx = 0; kx < gn[0]; kx++)
            for (int ky = 0; ky < gn[1]; ky++)
                for (int kz = 0; kz < gn[2]; kz++) {
                    size_t id = gidx(kx, ky, kz);
                    if (kx == 0 && ky == 0 && kz == 0) { grid[id] = 0; continue; }
                    int mx = kx < (gn[0] + 1) / 2 ? kx : kx - gn[0];
                    int my = ky < (gn[1] + 1) / 2 ? ky : ky - gn[1];
                    int mz = kz < (gn[2] + 1) / 2 ? kz : kz - gn[2];
                    double hx = mx * recip[0][0];
                    double hy = mx * recip[1][0] + my * recip[1][1];
                    double hz = mx * recip[2][0] + my * recip[2][1] + mz * recip[2][2];
                    double m2 = hx * hx + hy * hy + hz * hz;
                    double den = m2 * moduli[0][kx] * moduli[1][ky] * moduli[2][kz];
                    grid[id] *= sc * std::exp(-expf * m2) / den;
                }
    }

    // computeFixedPotentialFromGrid (:3368-3530) / computeInducedPotentialFromGrid (:3575-3737)
    void gather(std::vector<double>& out) {
        out.assign((size_t)35 * n, 0.0);
        for (int i = 0; i < n; i++) {
            double acc[35] = {0};
            for (int ix = 0; ix < kOrder; ix++) {
                int x = (igrid[3 * i] + ix) % gn[0];
                for (int iy = 0; iy < kOrder; iy++) {
                    int y = (igrid[3 * i + 1] + iy) % gn[1];
                    for (int iz = 0; iz < kOrder; iz++) {
                        int z = (igrid[3 * i + 2] + iz) % gn[2];
                        double g = grid[gidx(x, y, z)].real();
                        for (int k = 0; k < 35; k++)
                            acc[k] += g * th(0, i, ix, kPhiExp[k][0]) * th(1, i, iy, kPhiExp[k][1]) * th(2, i, iz, kPhiExp[k][2]);
                    }
                }
            }
            for (int k = 0; k < 35; k++) out[(size_t)35 * i + k] = acc[k];
        }
    }

    void recip_cycle(const std::vector<double>& m20, int ncomp, std::vector<double>& out) {
        spread(m20, ncomp);
        fft3d(grid, gn[0], gn[1], gn[2], -1);
        convolve();
        fft3d(grid, gn[0], gn[1], gn[2], +1);
        gather(out);
    }

    // E = -A * grad_frac(phi)  (recordFixedMultipoleField :4026-4037, recordInducedDipoleField :4046-4058)
    V3 field_from(const std::vector<double>& ph, int i) const {
        V3 e;
        for (int a = 0; a < 3; a++)
            e[a] = -(A[a][0] * ph[35 * i + 1] + A[a][1] * ph[35 * i + 2] + A[a][2] * ph[35 * i + 3]);
        return e;
    }

    // transformPotentialToCartesianCoordinates (:3171-3267): chain rule for
    // the Cartesian derivative tensors up to third order.
    void cartesian_potential(const std::vector<double>& ph, int i, double& c0, V3& c1, double c2[6], double c3[10]) const {
        const double* f = &ph[35 * (size_t)i];
        c0 = f[0];
        for (int a = 0; a < 3; a++) c1[a] = A[a][0] * f[1] + A[a][1] * f[2] + A[a][2] * f[3];
        for (int a = 0; a < 3; a++)
            for (int b = a; b < 3; b++) {
                double s = 0;
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) {
                        int e[3] = {0, 0, 0};
                        e[k]++; e[l]++;
                        s += A[a][k] * A[b][l] * f[phi_index(e[0], e[1], e[2])];
                    }
                c2[idx2(a, b)] = s;
            }
        for (int a = 0; a < 3; a++)
            for (int b = a; b < 3; b++)
                for (int c = b; c < 3; c++) {
                    double s = 0;
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++)
                            for (int o = 0; o < 3; o++) {
                                int e[3] = {0, 0, 0};
                                e[k]++; e[l]++; e[o]++;
                                s += A[a][k] * A[b][l] * A[c][o] * f[phi_index(e[0], e[1], e[2])];
                            }
                    c3[idx3(a, b, c)] = s;
                }
    }

    // Torque on the permanent moments (dipole d) in a potential with Cartesian
    // derivative tensors c1..c3: tau_k = -dE/dtheta_k with
    // E = d.c1 + Q:c2 + O:c3 (full contractions).  Same quantity as the hand
    // expansion at MPIDReferenceForce.cpp:3768-3799 and :3905-3946.
    V3 multipole_torque(const V3& d, const double* Q, const double* O, const V3& c1, const double* c2, const double* c3) const {
        V3 t;
        for (int k = 0; k < 3; k++) {
            double rq[6], ro[10];
            rot_quad(k, Q, rq);
            rot_oct(k, O, ro);
            t[k] = -(rot_vec(k, d).dot(c1) + contract2(rq, c2) + contract3(ro, c3));
        }
        return t;
    }

    // ---- real space -------------------------------------------------------
    // calculateFixedMultipoleFieldPairIxn (PME), :2812-2920 with
    // getDampedInverseDistances :2678-2718.
    void fixed_field_pair(const Pair& P, std::vector<V3>& E) const {
        const Site &I = site[P.i], &J = site[P.j];
        V3 dr = J.pos - I.pos;
        min_image(dr);
        double r2 = dr.dot(dr), r = std::sqrt(r2);
        double ar = S.alpha * r;
        double bn0 = std::erfc(ar) / r;
        double al2 = 2.0 * S.alpha * S.alpha;
        double aln = S.alpha > 0.0 ? 1.0 / (kSqrtPi * S.alpha) : 0.0;
        double ex = std::exp(-ar * ar);
        aln *= al2; double bn1 = (bn0 + aln * ex) / r2;
        aln *= al2; double bn2 = (3.0 * bn1 + aln * ex) / r2;
        aln *= al2; double bn3 = (5.0 * bn2 + aln * ex) / r2;
        aln *= al2; double bn4 = (7.0 * bn3 + aln * ex) / r2;
        double sf[4] = {1, 1, 1, 1};
        double damp = I.damp * J.damp;
        if (damp != 0.0) {
            double a = (P.pscale == 0.0) ? I.thole + J.thole : S.default_thole;
            double u = a * r / damp;
            if (u < 50.0) {
                double e = std::exp(-u), u2 = u * u, u3 = u2 * u, u4 = u3 * u, u5 = u4 * u;
                sf[0] = 1.0 - e * (1.0 + u + 0.5 * u2);
                sf[1] = 1.0 - e * (1.0 + u + 0.5 * u2 + u3 / 6.0);
                sf[2] = 1.0 - e * (1.0 + u + 0.5 * u2 + u3 / 6.0 + u4 / 30.0);
                sf[3] = 1.0 - e * (1.0 + u + 0.5 * u2 + u3 / 6.0 + 4.0 * u4 / 105.0 + u5 / 210.0);
            }
        }
        double ds = P.pscale;  // dScale == pScale
        double r3 = r * r2, r5 = r3 * r2, r7 = r5 * r2, r9 = r7 * r2;
        double d3 = (1.0 - ds * sf[0]) / r3, d5 = 3.0 * (1.0 - ds * sf[1]) / r5;
        double d7 = 15.0 * (1.0 - ds * sf[2]) / r7, d9 = 105.0 * (1.0 - ds * sf[3]) / r9;
        // multipole contractions with r
        auto contract = [&](const Site& s, V3& qr, double& qrr, V3& orr, double& orrr, double& dr_) {
            for (int a = 0; a < 3; a++) {
                qr[a] = 0;
                for (int b = 0; b < 3; b++) qr[a] += s.Q[idx2(a, b)] * dr[b];
            }
            qrr = qr.dot(dr);
            for (int a = 0; a < 3; a++) {
                orr[a] = 0;
                for (int b = 0; b < 3; b++)
                    for (int c = 0; c < 3; c++) orr[a] += s.O[idx3(a, b, c)] * dr[b] * dr[c];
            }
            orrr = orr.dot(dr);
            dr_ = s.d.dot(dr);
        };
        V3 qi, qj, oi, oj;
        double qir, qjr, oir, ojr, dir, djr;
        contract(I, qi, qir, oi, oir, dir);
        contract(J, qj, qjr, oj, ojr, djr);
        // field at i from j and at j from i, Ewald minus scaled bare
        auto fld = [&](double b1, double b2, double b3, double b4, bool atI) {
            if (atI)
                return qj * (2.0 * b2) - J.d * b1 - dr * (b1 * J.q - b2 * djr + b3 * qjr) - oj * (3.0 * b3) + dr * (b4 * ojr);
            return qi * (-2.0 * b2) - I.d * b1 + dr * (b1 * I.q + b2 * dir + b3 * qir) - oi * (3.0 * b3) + dr * (b4 * oir);
        };
        E[P.i] += fld(bn1, bn2, bn3, bn4, true) - fld(d3, d5, d7, d9, true);
        E[P.j] += fld(bn1, bn2, bn3, bn4, false) - fld(d3, d5, d7, d9, false);
    }

    // calculateDirectInducedDipolePairIxns (PME), :4161-4281
    void induced_field_pair(const Pair& P, const std::vector<V3>& m, std::vector<V3>& E) const {
        const Site &I = site[P.i], &J = site[P.j];
        V3 dr = J.pos - I.pos;
        min_image(dr);
        double r2 = dr.dot(dr), r = std::sqrt(r2);
        double ar = S.alpha * r;
        double bn0 = std::erfc(ar) / r;
        double al2 = 2.0 * S.alpha * S.alpha;
        double aln = S.alpha > 0.0 ? 1.0 / (kSqrtPi * S.alpha) : 0.0;
        double ex = std::exp(-ar * ar);
        aln *= al2; double bn1 = (bn0 + aln * ex) / r2;
        aln *= al2; double bn2 = (3.0 * bn1 + aln * ex) / r2;
        double s3 = 1.0, s5 = 1.0;
        double damp = I.damp * J.damp;
        if (damp != 0.0) {
            double a = (P.pscale == 0.0) ? I.thole + J.thole : S.default_thole;
            double u = a * r / damp;
            if (u < 50.0) {
                double e = std::exp(-u);
                s3 = 1.0 - e * (1.0 + u + 0.5 * u * u);
                s5 = 1.0 - e * (1.0 + u + 0.5 * u * u + u * u * u / 6.0);
            }
        }
        double r3 = r * r2, r5 = r3 * r2;
        double rr3 = (1.0 - s3) / r3, rr5 = 3.0 * (1.0 - s5) / r5;
        double f1 = rr3 - bn1, f2 = bn2 - rr5;
        E[P.i] += dr * (m[P.j].dot(dr) * f2) + m[P.j] * f1;
        E[P.j] += dr * (m[P.i].dot(dr) * f2) + m[P.i] * f1;
    }

    // Field gradient of the induced dipoles, the extra terms of
    // calculateDirectInducedDipolePairIxns (PME) for Extrapolated
    // polarization, :4225-4281: at i from mu_j with dr = r_j - r_i added, at
    // j from mu_i subtracted (components xx yy zz xy xz yz).
    void induced_grad_pair(const Pair& P, const std::vector<V3>& m, std::vector<double>& G) const {
        const Site &I = site[P.i], &J = site[P.j];
        V3 dr = J.pos - I.pos;
        min_image(dr);
        double r2 = dr.dot(dr), r = std::sqrt(r2);
        double ar = S.alpha * r;
        double bn0 = std::erfc(ar) / r;
        double al2 = 2.0 * S.alpha * S.alpha;
        double aln = S.alpha > 0.0 ? 1.0 / (kSqrtPi * S.alpha) : 0.0;
        double ex = std::exp(-ar * ar);
        aln *= al2; double bn1 = (bn0 + aln * ex) / r2;
        aln *= al2; double bn2 = (3.0 * bn1 + aln * ex) / r2;
        aln *= al2; double bn3 = (5.0 * bn2 + aln * ex) / r2;
        double s5 = 1.0, s7 = 1.0;
        double damp = I.damp * J.damp;
        if (damp != 0.0) {
            double a = (P.pscale == 0.0) ? I.thole + J.thole : S.default_thole;
            double u = a * r / damp;
            if (u < 50.0) {
                double e = std::exp(-u), u2 = u * u, u3 = u2 * u, u4 = u3 * u;
                s5 = 1.0 - e * (1.0 + u + 0.5 * u2 + u3 / 6.0);
                s7 = 1.0 - e * (1.0 + u + 0.5 * u2 + u3 / 6.0 + u4 / 30.0);
            }
        }
        double r5 = r2 * r2 * r, r7 = r5 * r2;
        double f2 = bn2 - 3.0 * (1.0 - s5) / r5, f3 = bn3 - 15.0 * (1.0 - s7) / r7;
        // the gradient tensor of the field of dipole u at separation dr
        auto grad = [&](const V3& u, double g[6]) {
            const int ab[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
            double ud = u.dot(dr);
            for (int k = 0; k < 6; k++) {
                int a = ab[k][0], b = ab[k][1];
                g[k] = ud * dr[a] * dr[b] * f3 - (u[a] * dr[b] + u[b] * dr[a] + (a == b ? ud : 0.0)) * f2;
            }
        };
        double gi[6], gj[6];
        grad(m[P.j], gi);
        grad(m[P.i], gj);
        for (int k = 0; k < 6; k++) {
            G[6 * (size_t)P.i + k] += gi[k];
            G[6 * (size_t)P.j + k] -= gj[k];
        }
    }

    // ---- fixed field: calculateFixedMultipoleField (PME), :2922-2949
    void fixed_field() {
        if (periodic) {
            fractional_multipoles();
            recip_cycle(mfrac, 20, phi);
        } else {
            mfrac.assign(20 * (size_t)n, 0.0);
            phi.assign(35 * (size_t)n, 0.0);
        }
        E0.assign(n, V3());
        for (int i = 0; i < n; i++) E0[i] = field_from(phi, i);
        double term = (4.0 / 3.0) * (S.alpha * S.alpha * S.alpha) / kSqrtPi;
        for (int i = 0; i < n; i++) E0[i] += site[i].d * term;
        for (const Pair& p : pairs) fixed_field_pair(p, E0);
    }

    // induced dipole field T*mu: real + reciprocal + self (:4073-4141)
    void induced_field(const std::vector<V3>& m, std::vector<V3>& E, std::vector<double>* phd = nullptr) {
        E.assign(n, V3());
        for (const Pair& p : pairs) induced_field_pair(p, m, E);
        std::vector<double> m20((size_t)20 * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < 3; j++) {
                double v = 0;
                for (int k = 0; k < 3; k++) v += A[k][j] * m[i][k];
                m20[20 * (size_t)i + 1 + j] = v;
            }
        std::vector<double> ph;
        if (periodic) {
            recip_cycle(m20, 4, ph);
            for (int i = 0; i < n; i++) E[i] += field_from(ph, i);
        } else {
            ph.assign(35 * (size_t)n, 0.0);
        }
        double term = (4.0 / 3.0) * (S.alpha * S.alpha * S.alpha) / kSqrtPi;
        for (int i = 0; i < n; i++) E[i] += m[i] * term;
        if (phd) *phd = ph;
    }

    // Field (real + reciprocal + self) and field gradient (real + reciprocal,
    // no self term) of the dipoles m: calculateInducedDipoleFields (PME)
    // with the gradient terms of Extrapolated polarization, :4073-4141.  The
    // reciprocal gradient is minus the Cartesian second derivatives of the
    // dipoles' reciprocal potential, A phi_frac A^T (:4094-4127).
    void induced_field_grad(const std::vector<V3>& m, std::vector<V3>& E, std::vector<double>& G) {
        std::vector<double> ph;
        induced_field(m, E, &ph);
        G.assign(6 * (size_t)n, 0.0);
        for (const Pair& p : pairs) induced_grad_pair(p, m, G);
        const int ab[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 6; c++) {
                const int a = ab[c][0], b = ab[c][1];
                double v = 0;
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) {
                        int e[3] = {0, 0, 0};
                        e[k]++;
                        e[l]++;
                        v += A[a][k] * A[b][l] * ph[35 * (size_t)i + phi_index(e[0], e[1], e[2])];
                    }
                G[6 * (size_t)i + c] -= v;
            }
    }

    // convergeInduceDipolesByExtrapolation, :1125-1180, with the partial sums
    // of setExtrapolationCoefficients, :173-183: mu_(0) = alpha E0,
    // mu_(m+1) = alpha E(mu_(m)), mu = sum_m c'_m mu_(m), c'_m = sum_{j>=m} c_j.
    void extrapolate() {
        const int N = S.n_coeff;
        cpart.assign(N, 0.0);
        for (int i = 0; i < N; i++)
            for (int j = i; j < N; j++) cpart[i] += S.coeffs[j];
        pt_mu.assign(1, mu);
        pt_E.clear();
        pt_G.clear();
        for (int order = 1; order < N; order++) {
            std::vector<V3> E;
            std::vector<double> G;
            induced_field_grad(pt_mu.back(), E, G);
            std::vector<V3> next(n);
            for (int i = 0; i < n; i++) next[i] = apply_pol(i, E[i]);
            pt_E.push_back(E);
            pt_G.push_back(G);
            pt_mu.push_back(next);
        }
        mu.assign(n, V3());
        for (int o = 0; o < N; o++)
            for (int i = 0; i < n; i++) mu[i] += pt_mu[o][i] * cpart[o];
        converged = true;
        iterations = 0;
        last_eps = 0;
    }

    // The dipole-response forces and torques of Extrapolated polarization,
    // :4955-4985: sum over l, m (l + m <= N - 2) of c'_(l+m+1) times mu_(l)
    // contracted with the field gradient of mu_(m) (force) or crossed with
    // its field (torque, anisotropic sites only).  Terms with
    // |c'| < 1e-6 are skipped as in the reference.
    void extrapolation_response(std::vector<V3>& F, std::vector<V3>& T) const {
        const int N = (int)cpart.size();
        for (int i = 0; i < n; i++)
            for (int l = 0; l < N - 1; l++)
                for (int m = 0; m < N - 1 - l; m++) {
                    const double p = cpart[l + m + 1];
                    if (std::fabs(p) < 1e-6) continue;
                    const V3& u = pt_mu[l][i];
                    const double* g = &pt_G[m][6 * (size_t)i];
                    const double pf = p * kElectric;
                    F[i][0] += pf * (u[0] * g[0] + u[1] * g[3] + u[2] * g[4]);
                    F[i][1] += pf * (u[0] * g[3] + u[1] * g[1] + u[2] * g[5]);
                    F[i][2] += pf * (u[0] * g[4] + u[1] * g[5] + u[2] * g[2]);
                    if (site[i].aniso) T[i] += u.cross(pt_E[m][i]) * pf;
                }
    }

    V3 apply_pol(int i, const V3& e) const {
        const double* p = site[i].labpol;
        return V3(p[0] * e[0] + p[1] * e[1] + p[2] * e[2], p[1] * e[0] + p[3] * e[1] + p[4] * e[2],
                  p[2] * e[0] + p[4] * e[1] + p[5] * e[2]);
    }

    // DIIS coefficients (:1254-1291) through the pseudo-inverse of B.
    static void diis_coefficients(const std::vector<std::vector<V3>>& errs, std::vector<double>& c) {
        int m = (int)errs.size();
        c.assign(m, 0.0);
        if (m == 1) { c[0] = 1.0; return; }
        int R = m + 1;
        std::vector<double> B(R * R, 0.0);
        for (int i = 0; i < m; i++) B[(i + 1) * R] = B[i + 1] = -1.0;
        for (int i = 0; i < m; i++)
            for (int j = i; j < m; j++) {
                double s = 0;
                for (size_t k = 0; k < errs[i].size(); k++) s += errs[i][k].dot(errs[j][k]);
                B[(i + 1) * R + j + 1] = B[(j + 1) * R + i + 1] = s;
            }
        // cyclic Jacobi eigen-decomposition B = V diag(w) V^T
        std::vector<double> V(R * R, 0.0), a = B;
        for (int i = 0; i < R; i++) V[i * R + i] = 1.0;
        for (int sweep = 0; sweep < 100; sweep++) {
            double off = 0;
            for (int p = 0; p < R; p++)
                for (int q = p + 1; q < R; q++) off += a[p * R + q] * a[p * R + q];
            if (off < 1e-300) break;
            for (int p = 0; p < R; p++)
                for (int q = p + 1; q < R; q++) {
                    double apq = a[p * R + q];
                    if (std::fabs(apq) < 1e-300) continue;
                    double th = (a[q * R + q] - a[p * R + p]) / (2.0 * apq);
                    double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                    double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                    for (int k = 0; k < R; k++) {
                        double akp = a[k * R + p], akq = a[k * R + q];
                        a[k * R + p] = cs * akp - sn * akq;
                        a[k * R + q] = sn * akp + cs * akq;
                    }
                    for (int k = 0; k < R; k++) {
                        double apk = a[p * R + k], aqk = a[q * R + k];
                        a[p * R + k] = cs * apk - sn * aqk;
                        a[q * R + k] = sn * apk + cs * aqk;
                    }
                    for (int k = 0; k < R; k++) {
                        double vkp = V[k * R + p], vkq = V[k * R + q];
                        V[k * R + p] = cs * vkp - sn * vkq;
                        V[k * R + q] = sn * vkp + cs * vkq;
                    }
                }
        }
        double smax = 0;
        for (int i = 0; i < R; i++) smax = std::max(smax, std::fabs(a[i * R + i]));
        double tol = R * smax * 2.220446049250313e-16;  // JAMA::SVD::rank tolerance
        for (int i = 1; i < R; i++) {
            double d = 0;
            for (int j = 0; j < R; j++) {
                double w = a[j * R + j];
                if (std::fabs(w) > tol) d -= V[0 * R + j] * V[i * R + j] / w;
            }
            c[i - 1] = d;
        }
    }

    // calculateInducedDipoles (:1293-1329) + convergeInduceDipolesByDIIS (:1182-1252)
    void solve() {
        mu.assign(n, V3());
        std::vector<V3> pE(n);
        for (int i = 0; i < n; i++) pE[i] = apply_pol(i, E0[i]);
        mu = pE;
        if (S.polarization == 1) { converged = true; iterations = 0; return; }
        if (S.polarization == 2) { extrapolate(); return; }
        std::vector<std::vector<V3>> prevD, prevE;
        const int maxPrev = 20;
        for (int it = 0;; it++) {
            std::vector<V3> Ei;
            induced_field(mu, Ei);
            std::vector<V3> nd(n), er(n);
            double e2 = 0;
            for (int i = 0; i < n; i++) {
                nd[i] = pE[i] + apply_pol(i, Ei[i]);
                er[i] = nd[i] - mu[i];
                e2 += er[i].dot(er[i]);
            }
            prevD.push_back(nd);
            prevE.push_back(er);
            double eps = kDebye * std::sqrt(e2 / n);
            last_eps = eps;
            // progress of long solves (the benchmark's timing of large boxes)
            if (progress_lines()) {
                std::fprintf(stderr, "[oracle] DIIS iteration %d: eps %.3e Debye\n", it, eps);
                std::fflush(stderr);
            }
            if (eps < S.target_eps) converged = true;
            if (eps < S.target_eps || it == S.max_iter) { iterations = it; return; }
            if ((int)prevE.size() > maxPrev) {
                prevE.erase(prevE.begin());
                prevD.erase(prevD.begin());
            }
            std::vector<double> c;
            diis_coefficients(prevE, c);
            for (int i = 0; i < n; i++) {
                V3 v;
                for (size_t j = 0; j < c.size(); j++) v += prevD[j][i] * c[j];
                mu[i] = v;
            }
        }
    }

    // ---- calculatePmeDirectElectrostaticPairIxn, :4335-4920
    double electro_pair(const Pair& P, std::vector<V3>& F, std::vector<V3>& T) const {
        const Site &I = site[P.i], &J = site[P.j];
        V3 dr = J.pos - I.pos;
        min_image(dr);
        double r2 = dr.dot(dr), r = std::sqrt(r2);
        // QI frame: z along r, x from z plus a lab axis (formQIRotationMatrix :650-681)
        V3 ez = dr * (1.0 / r);
        V3 ex = ez;
        if (I.pos[1] != J.pos[1] || I.pos[2] != J.pos[2]) ex[0] += 1.0; else ex[1] += 1.0;
        ex -= ez * ez.dot(ex);
        normalize(ex);
        V3 ey = ez.cross(ex);
        V3 fr[3] = {ex, ey, ez};
        // QI-frame spherical moments: rotate the Cartesian tensors, then project.
        double qa[16], qb[16];
        auto qi_moments = [&](const Site& s, double* out) {
            V3 d = frame_vec(fr, s.d);
            double Q[6], O[10];
            frame_quad(fr, s.Q, Q);
            frame_oct(fr, s.O, O);
            out[0] = s.q;
            sph_dip(d, out + 1);
            sph_quad(Q, out + 4);
            sph_oct(O, out + 9);
        };
        qi_moments(I, qa);
        qi_moments(J, qb);
        // rotational derivatives of the QI moments about the QI axes
        double qaX[3][16], qbX[3][16];
        auto qi_rot = [&](const Site& s, double out[3][16]) {
            V3 d = frame_vec(fr, s.d);
            double Q[6], O[10];
            frame_quad(fr, s.Q, Q);
            frame_oct(fr, s.O, O);
            for (int k = 0; k < 3; k++) {
                double rq[6], ro[10];
                V3 rd = rot_vec(k, d);
                rot_quad(k, Q, rq);
                rot_oct(k, O, ro);
                out[k][0] = 0.0;
                sph_dip(rd, out[k] + 1);
                sph_quad(rq, out[k] + 4);
                sph_oct(ro, out[k] + 9);
            }
        };
        qi_rot(I, qaX);
        qi_rot(J, qbX);
        // induced dipoles in the QI frame, spherical order (z, x, y), halved
        double ua[3], ub[3];
        {
            V3 a = frame_vec(fr, mu[P.i]), b = frame_vec(fr, mu[P.j]);
            ua[0] = 0.5 * a[2]; ua[1] = 0.5 * a[0]; ua[2] = 0.5 * a[1];
            ub[0] = 0.5 * b[2]; ub[1] = 0.5 * b[0]; ub[2] = 0.5 * b[1];
        }
        double Va[16] = {0}, Vb[16] = {0}, VaR[16] = {0}, VbR[16] = {0}, Vad[3] = {0}, Vbd[3] = {0};
        double ri[9], ar[10], b[6];
        double pre = kElectric;
        double rInv = 1.0 / r;
        ri[1] = pre * rInv;
        for (int i = 2; i < 9; i++) ri[i] = ri[i - 1] * rInv;
        ar[1] = S.alpha * r;
        for (int i = 2; i < 10; i++) ar[i] = ar[i - 1] * ar[1];
        double X = 2.0 * std::exp(-ar[2]) / kSqrtPi;
        double ms = P.mscale, ps = P.pscale, ds = ps, us = 1.0;
        {
            int dfac = 1, fc = 1;
            double tmp = ar[1];
            b[1] = -std::erf(ar[1]);
            for (int i = 2; i < 6; i++) {
                b[i] = b[i - 1] + tmp * X / dfac;
                fc += 2;
                dfac *= fc;
                tmp *= 2.0 * ar[2];
            }
        }
        double dmp = I.damp * J.damp;
        double a = (ps == 0.0) ? I.thole + J.thole : S.default_thole;
        double u = std::fabs(dmp) > 1.0e-5 ? r / dmp : 1e10;
        double au = a * u;
        double e = au < 50.0 ? std::exp(-au) : 0.0;
        double au2 = au * au, au3 = au2 * au, au4 = au3 * au, au5 = au4 * au, au6 = au5 * au;
        // Thole factors (:4509-4523)
        double th_c = 1.0 - e * (1.0 + au + 0.5 * au2);
        double th_d0 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 4.0);
        double th_d1 = 1.0 - e * (1.0 + au + 0.5 * au2);
        double th_q0 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 18.0);
        double th_q1 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0);
        double th_o0 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 24.0 + au5 / 120.0);
        double th_o1 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 30.0);
        double dth_c = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 4.0);
        double dth_d0 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 12.0);
        double dth_d1 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0);
        double dth_q0 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 24.0 + au5 / 72.0);
        double dth_q1 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 24.0);
        double dth_o0 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 24.0 + au5 / 120.0 + au6 / 600.0);
        double dth_o1 = 1.0 - e * (1.0 + au + 0.5 * au2 + au3 / 6.0 + au4 / 25.0 + au5 / 150.0);
        const double s3 = std::sqrt(3.0), s5 = std::sqrt(5.0), s6 = std::sqrt(6.0), s2 = std::sqrt(2.0);
        double eP, dP, eU, dU;

        // C-C (m=0)
        eP = ri[1] * (ms + b[2] - ar[1] * X);
        dP = -0.5 * (ms + b[2]) * ri[2];
        Va[0] = eP * qb[0]; Vb[0] = eP * qa[0]; VaR[0] = dP * qb[0]; VbR[0] = dP * qa[0];
        // C-D, C-Uind (m=0)
        eP = ri[2] * (ms + b[2]);
        eU = 2.0 * ri[2] * (ps * th_c + b[2]);
        dP = -ri[3] * (ms + b[2] + ar[3] * X);
        dU = -4.0 * ri[3] * (ds * dth_c + b[2] + ar[3] * X);
        Va[0] += -(eP * qb[1] + eU * ub[0]); Vb[1] = -(eP * qa[0]);
        VaR[0] += -(dP * qb[1] + dU * ub[0]); VbR[1] = -(dP * qa[0]);
        Vbd[0] = -(eU * qa[0]);
        // D-C, Uind-C (m=0)
        Va[1] = eP * qb[0]; Vb[0] += eP * qa[1] + eU * ua[0];
        VaR[1] = dP * qb[0]; VbR[0] += dP * qa[1] + dU * ua[0];
        Vad[0] = eU * qb[0];
        // D-D, D-Uind (m=0)
        eP = -(2.0 / 3.0) * ri[3] * (3.0 * (ms + b[3]) + ar[3] * X);
        eU = -2.0 * (2.0 / 3.0) * ri[3] * (3.0 * (ds * th_d0 + b[3]) + ar[3] * X);
        dP = ri[4] * (3.0 * (ms + b[3]) + 2.0 * ar[5] * X);
        dU = 2.0 * ri[4] * (6.0 * (ds * dth_d0 + b[3]) + 4.0 * ar[5] * X);
        Va[1] += eP * qb[1] + eU * ub[0]; Vb[1] += eP * qa[1] + eU * ua[0];
        VaR[1] += dP * qb[1] + dU * ub[0]; VbR[1] += dP * qa[1] + dU * ua[0];
        Vad[0] += eU * qb[1]; Vbd[0] += eU * qa[1];
        // D-D, D-Uind (m=1)
        eP = ri[3] * (ms + b[3] - (2.0 / 3.0) * ar[3] * X);
        eU = 2.0 * ri[3] * (ds * th_d1 + b[3] - (2.0 / 3.0) * ar[3] * X);
        dP = -1.5 * ri[4] * (ms + b[3]);
        dU = -6.0 * ri[4] * (ds * dth_d1 + b[3]);
        for (int c = 0; c < 2; c++) {
            Va[2 + c] = eP * qb[2 + c] + eU * ub[1 + c]; Vb[2 + c] = eP * qa[2 + c] + eU * ua[1 + c];
            VaR[2 + c] = dP * qb[2 + c] + dU * ub[1 + c]; VbR[2 + c] = dP * qa[2 + c] + dU * ua[1 + c];
            Vad[1 + c] = eU * qb[2 + c]; Vbd[1 + c] = eU * qa[2 + c];
        }
        // C-Q, Q-C (m=0)
        eP = (ms + b[3]) * ri[3];
        dP = -(1.0 / 3.0) * ri[4] * (4.5 * (ms + b[3]) + 2.0 * ar[5] * X);
        Va[0] += eP * qb[4]; Vb[4] = eP * qa[0]; VaR[0] += dP * qb[4]; VbR[4] = dP * qa[0];
        Va[4] = eP * qb[0]; Vb[0] += eP * qa[4]; VaR[4] = dP * qb[0]; VbR[0] += dP * qa[4];
        // D-Q, Uind-Q (m=0)
        eP = ri[4] * (3.0 * (ms + b[3]) + (4.0 / 3.0) * ar[5] * X);
        eU = 2.0 * ri[4] * (3.0 * (ds * th_q0 + b[3]) + (4.0 / 3.0) * ar[5] * X);
        dP = -(4.0 / 3.0) * ri[5] * (4.5 * (ms + b[3]) + (1.0 + ar[2]) * ar[5] * X);
        dU = -2.0 * (4.0 / 3.0) * ri[5] * (9.0 * (ds * dth_q0 + b[3]) + 2.0 * (1.0 + ar[2]) * ar[5] * X);
        Va[1] += eP * qb[4]; Vb[4] += eP * qa[1] + eU * ua[0];
        VaR[1] += dP * qb[4]; VbR[4] += dP * qa[1] + dU * ua[0];
        Vad[0] += eU * qb[4];
        // Q-D, Q-Uind (m=0)
        Va[4] += -(eP * qb[1] + eU * ub[0]); Vb[1] += -(eP * qa[4]);
        VaR[4] += -(dP * qb[1] + dU * ub[0]); VbR[1] += -(dP * qa[4]);
        Vbd[0] += -(eU * qa[4]);
        // D-Q, Uind-Q (m=1)
        eP = -s3 * ri[4] * (ms + b[3]);
        eU = -2.0 * s3 * ri[4] * (ps * th_q1 + b[3]);
        dP = (4.0 / s3) * ri[5] * (1.5 * (ms + b[3]) + 0.5 * ar[5] * X);
        dU = 2.0 * (4.0 / s3) * ri[5] * (3.0 * (ds * dth_q1 + b[3]) + ar[5] * X);
        for (int c = 0; c < 2; c++) {
            Va[2 + c] += eP * qb[5 + c]; Vb[5 + c] = eP * qa[2 + c] + eU * ua[1 + c];
            VaR[2 + c] += dP * qb[5 + c]; VbR[5 + c] = dP * qa[2 + c] + dU * ua[1 + c];
            Vad[1 + c] += eU * qb[5 + c];
        }
        // Q-D, Q-Uind (m=1)
        for (int c = 0; c < 2; c++) {
            Va[5 + c] = -(eP * qb[2 + c] + eU * ub[1 + c]); Vb[2 + c] += -(eP * qa[5 + c]);
            VaR[5 + c] = -(dP * qb[2 + c] + dU * ub[1 + c]); VbR[2 + c] += -(dP * qa[5 + c]);
            Vbd[1 + c] += -(eU * qa[5 + c]);
        }
        // Q-Q (m=0)
        eP = ri[5] * (6.0 * (ms + b[4]) + (4.0 / 45.0) * (-3.0 + 10.0 * ar[2]) * ar[5] * X);
        dP = -(1.0 / 9.0) * ri[6] * (135.0 * (ms + b[4]) + 4.0 * (1.0 + 2.0 * ar[2]) * ar[7] * X);
        Va[4] += eP * qb[4]; Vb[4] += eP * qa[4]; VaR[4] += dP * qb[4]; VbR[4] += dP * qa[4];
        // Q-Q (m=1)
        eP = -(4.0 / 15.0) * ri[5] * (15.0 * (ms + b[4]) + ar[5] * X);
        dP = ri[6] * (10.0 * (ms + b[4]) + (4.0 / 3.0) * ar[7] * X);
        for (int c = 5; c < 7; c++) {
            Va[c] += eP * qb[c]; Vb[c] += eP * qa[c]; VaR[c] += dP * qb[c]; VbR[c] += dP * qa[c];
        }
        // Q-Q (m=2)
        eP = ri[5] * (ms + b[4] - (4.0 / 15.0) * ar[5] * X);
        dP = -2.5 * (ms + b[4]) * ri[6];
        for (int c = 7; c < 9; c++) {
            Va[c] = eP * qb[c]; Vb[c] = eP * qa[c]; VaR[c] = dP * qb[c]; VbR[c] = dP * qa[c];
        }
        // C-O, O-C (m=0)
        eP = ri[4] * (-ms - b[3] - (4.0 / 15.0) * ar[5] * X);
        dP = 0.5 * (4.0 / 15.0) * ri[5] * (15.0 * (ms + b[3]) + 2.0 * (2.0 * ar[5] + ar[7]) * X);
        Va[0] += eP * qb[9]; Vb[9] = eP * qa[0]; VaR[0] += dP * qb[9]; VbR[9] = dP * qa[0];
        Va[9] = -eP * qb[0]; Vb[0] -= eP * qa[9]; VaR[9] = -dP * qb[0]; VbR[0] -= dP * qa[9];
        // D-O, Uind-O (m=0)
        eP = -4.0 * ri[5] * (ms + b[4] + (2.0 / 15.0) * ar[7] * X);
        eU = -8.0 * ri[5] * (ds * th_o0 + b[4] + (2.0 / 15.0) * ar[7] * X);
        dP = 0.5 * (4.0 / 15.0) * ri[6] * (75.0 * (ms + b[4]) + 4.0 * (1.0 + ar[2]) * ar[7] * X);
        dU = 2.0 * (4.0 / 15.0) * ri[6] * (75.0 * (ds * dth_o0 + b[4]) + 4.0 * (1.0 + ar[2]) * ar[7] * X);
        Va[1] += eP * qb[9]; Vb[9] += eP * qa[1] + eU * ua[0];
        VaR[1] += dP * qb[9]; VbR[9] += dP * qa[1] + dU * ua[0];
        Va[9] += eP * qb[1] + eU * ub[0]; Vb[1] += eP * qa[9];
        VaR[9] += dP * qb[1] + dU * ub[0]; VbR[1] += dP * qa[9];
        Vad[0] += eU * qb[9]; Vbd[0] += eU * qa[9];
        // D-O, Uind-O (m=1)
        eP = s6 * (ms + b[4]) * ri[5];
        eU = 2.0 * s6 * (ds * th_o1 + b[4]) * ri[5];
        dP = -0.5 * 0.2 * std::sqrt(2.0 / 3.0) * ri[6] * (75.0 * (ms + b[4]) + 8.0 * ar[7] * X);
        dU = -2.0 * 0.2 * std::sqrt(2.0 / 3.0) * ri[6] * (75.0 * (ds * dth_o1 + b[4]) + 8.0 * ar[7] * X);
        for (int c = 0; c < 2; c++) {
            Va[2 + c] += eP * qb[10 + c]; Vb[10 + c] = eP * qa[2 + c] + eU * ua[1 + c];
            VaR[2 + c] += dP * qb[10 + c]; VbR[10 + c] = dP * qa[2 + c] + dU * ua[1 + c];
            Vad[1 + c] += eU * qb[10 + c];
        }
        for (int c = 0; c < 2; c++) {
            Va[10 + c] = eP * qb[2 + c] + eU * ub[1 + c]; Vb[2 + c] += eP * qa[10 + c];
            VaR[10 + c] = dP * qb[2 + c] + dU * ub[1 + c]; VbR[2 + c] += dP * qa[10 + c];
            Vbd[1 + c] += eU * qa[10 + c];
        }
        // Q-O, O-Q (m=0)
        eP = ri[6] * (-10.0 * (ms + b[4]) - (8.0 / 45.0) * (3.0 + 2.0 * ar[2]) * ar[7] * X);
        dP = 0.5 * (4.0 / 45.0) * ri[7] * (675.0 * (ms + b[4]) + 2.0 * (27.0 + 4.0 * ar[4]) * ar[7] * X);
        Va[4] += eP * qb[9]; Vb[9] += eP * qa[4]; VaR[4] += dP * qb[9]; VbR[9] += dP * qa[4];
        Va[9] -= eP * qb[4]; Vb[4] -= eP * qa[9]; VaR[9] -= dP * qb[4]; VbR[4] -= dP * qa[9];
        // Q-O, O-Q (m=1)
        eP = 5.0 * s2 * ri[6] * (ms + b[4] + (8.0 / 75.0) * ar[7] * X);
        dP = -0.5 * (std::sqrt(8.0) / 15.0) * ri[7] * (225.0 * (ms + b[4]) + 8.0 * (2.0 + ar[2]) * ar[7] * X);
        for (int c = 0; c < 2; c++) {
            Va[5 + c] += eP * qb[10 + c]; Vb[10 + c] += eP * qa[5 + c];
            VaR[5 + c] += dP * qb[10 + c]; VbR[10 + c] += dP * qa[5 + c];
            Va[10 + c] -= eP * qb[5 + c]; Vb[5 + c] -= eP * qa[10 + c];
            VaR[10 + c] -= dP * qb[5 + c]; VbR[5 + c] -= dP * qa[10 + c];
        }
        // Q-O, O-Q (m=2)
        eP = -s5 * (ms + b[4]) * ri[6];
        dP = 0.5 * (2.0 * s5 / 15.0) * ri[7] * (45.0 * (ms + b[4]) + 4.0 * ar[7] * X);
        for (int c = 0; c < 2; c++) {
            Va[7 + c] += eP * qb[12 + c]; Vb[12 + c] = eP * qa[7 + c];
            VaR[7 + c] += dP * qb[12 + c]; VbR[12 + c] = dP * qa[7 + c];
            Va[12 + c] = -eP * qb[7 + c]; Vb[7 + c] -= eP * qa[12 + c];
            VaR[12 + c] = -dP * qb[7 + c]; VbR[7 + c] -= dP * qa[12 + c];
        }
        // O-O (m=0..3)
        eP = ri[7] * (-20.0 * (ms + b[5]) - (8.0 / 1575.0) * (15.0 + 28.0 * ar[2] + 28.0 * ar[4]) * ar[7] * X);
        dP = 0.5 * (4.0 / 225.0) * ri[8] * (7875.0 * (ms + b[5]) + 4.0 * (41.0 - 4.0 * ar[2] + 4.0 * ar[4]) * ar[9] * X);
        Va[9] += eP * qb[9]; Vb[9] += eP * qa[9]; VaR[9] += dP * qb[9]; VbR[9] += dP * qa[9];
        eP = ri[7] * (15.0 * (ms + b[5]) + (8.0 / 525.0) * (-5.0 + 28.0 * ar[2]) * ar[7] * X);
        dP = -0.5 * (2.0 / 150.0) * ri[8] * (7875.0 * (ms + b[5]) + 32.0 * (3.0 + 2.0 * ar[2]) * ar[9] * X);
        for (int c = 10; c < 12; c++) {
            Va[c] += eP * qb[c]; Vb[c] += eP * qa[c]; VaR[c] += dP * qb[c]; VbR[c] += dP * qa[c];
        }
        eP = ri[7] * (-6.0 * (ms + b[5]) - (8.0 / 105.0) * ar[7] * X);
        dP = 0.5 * ri[8] * (42.0 * (ms + b[5]) + (16.0 / 15.0) * ar[9] * X);
        for (int c = 12; c < 14; c++) {
            Va[c] += eP * qb[c]; Vb[c] += eP * qa[c]; VaR[c] += dP * qb[c]; VbR[c] += dP * qa[c];
        }
        eP = ri[7] * ((ms + b[5]) - (8.0 / 105.0) * ar[7] * X);
        dP = -0.5 * 7.0 * (ms + b[5]) * ri[8];
        for (int c = 14; c < 16; c++) {
            Va[c] = eP * qb[c]; Vb[c] = eP * qa[c]; VaR[c] = dP * qb[c]; VbR[c] = dP * qa[c];
        }

        double energy = 0, fIZ = 0, fJZ = 0, EI[3] = {0, 0, 0}, EJ[3] = {0, 0, 0};
        for (int i = 0; i < 16; i++) {
            energy += 0.5 * (qa[i] * Va[i] + qb[i] * Vb[i]);
            fIZ += qa[i] * VaR[i];
            fJZ += qb[i] * VbR[i];
            for (int k = 0; k < 3; k++) {
                EI[k] += qaX[k][i] * Va[i];
                EJ[k] += qbX[k][i] * Vb[i];
            }
        }
        // induced torque intermediates (:4795-4816): rotational derivative of
        // the QI induced dipole dotted with the permanent-only field.
        double iEI[3], iEJ[3];
        iEI[0] = ua[2] * Vad[0] - ua[0] * Vad[2];
        iEJ[0] = ub[2] * Vbd[0] - ub[0] * Vbd[2];
        iEI[1] = ua[0] * Vad[1] - ua[1] * Vad[0];
        iEJ[1] = ub[0] * Vbd[1] - ub[1] * Vbd[0];
        iEI[2] = ua[1] * Vad[2] - ua[2] * Vad[1];
        iEJ[2] = ub[1] * Vbd[2] - ub[2] * Vbd[1];
        if (S.polarization == 0) {
            double eC = -2.0 * (4.0 / 3.0) * ri[3] * (3.0 * (us * th_d0 + b[3]) + ar[3] * X);
            double dC = 2.0 * ri[4] * (6.0 * (us * dth_d0 + b[3]) + 4.0 * ar[5] * X);
            iEI[0] += eC * ua[2] * ub[0];
            iEJ[0] += eC * ub[2] * ua[0];
            iEI[1] -= eC * ua[1] * ub[0];
            iEJ[1] -= eC * ub[1] * ua[0];
            fIZ += dC * ua[0] * ub[0];
            fJZ += dC * ub[0] * ua[0];
            eC = 4.0 * ri[3] * (us * th_d1 + b[3] - (2.0 / 3.0) * ar[3] * X);
            dC = -6.0 * ri[4] * (us * dth_d1 + b[3]);
            iEI[0] -= eC * ua[0] * ub[2];
            iEJ[0] -= eC * ub[0] * ua[2];
            iEI[1] += eC * ua[0] * ub[1];
            iEJ[1] += eC * ub[0] * ua[1];
            iEI[2] += eC * ua[1] * ub[2];
            iEJ[2] += eC * ub[1] * ua[2];
            if (S.flags & 1) {
                // the reference (:4854-4855) keeps only one of the two terms of
                // this rotational derivative; flag 1 adds the other one
                iEI[2] -= eC * ua[2] * ub[1];
                iEJ[2] -= eC * ub[2] * ua[1];
            }
            fIZ += dC * (ua[1] * ub[1] + ua[2] * ub[2]);
            fJZ += dC * (ub[1] * ua[1] + ub[2] * ua[2]);
        }
        double qf[3] = {rInv * (EI[1] + EJ[1] + iEI[1] + iEJ[1]), -rInv * (EI[0] + EJ[0] + iEI[0] + iEJ[0]), -(fJZ + fIZ)};
        double tI[3] = {-EI[0], -EI[1], -EI[2]}, tJ[3] = {-EJ[0], -EJ[1], -EJ[2]};
        if (I.aniso)
            for (int k = 0; k < 3; k++) tI[k] -= iEI[k];
        if (J.aniso)
            for (int k = 0; k < 3; k++) tJ[k] -= iEJ[k];
        V3 f = fr[0] * qf[0] + fr[1] * qf[1] + fr[2] * qf[2];
        V3 ti = fr[0] * tI[0] + fr[1] * tI[1] + fr[2] * tI[2];
        V3 tj = fr[0] * tJ[0] + fr[1] * tJ[1] + fr[2] * tJ[2];
        T[P.i] += ti;
        T[P.j] += tj;
        F[P.i] -= f;
        F[P.j] += f;
        return energy;
    }

    // computeReciprocalSpaceFixedMultipoleForceAndEnergy, :3739-3866
    double recip_fixed(std::vector<V3>& F, std::vector<V3>& T) const {
        double energy = 0;
        int ce[20][3];
        for (int k = 0; k < 20; k++) comp_exp(k, ce[k]);
        for (int i = 0; i < n; i++) {
            const Site& s = site[i];
            double c0, c2[6], c3[10];
            V3 c1;
            cartesian_potential(phi, i, c0, c1, c2, c3);
            V3 d = s.d;
            if (s.aniso) d += mu[i];
            T[i] += multipole_torque(d, s.Q, s.O, c1, c2, c3) * kElectric;
            const double* m = &mfrac[20 * (size_t)i];
            V3 f;
            for (int k = 0; k < 20; k++) {
                energy += m[k] * phi[35 * (size_t)i + phi_index(ce[k][0], ce[k][1], ce[k][2])];
                for (int a = 0; a < 3; a++) {
                    int e2[3] = {ce[k][0], ce[k][1], ce[k][2]};
                    e2[a]++;
                    f[a] += m[k] * phi[35 * (size_t)i + phi_index(e2[0], e2[1], e2[2])];
                }
            }
            f = f * kElectric;
            for (int a = 0; a < 3; a++) F[i][a] -= A[a][0] * f[0] + A[a][1] * f[1] + A[a][2] * f[2];
        }
        return 0.5 * kElectric * energy;
    }

    // computeReciprocalSpaceInducedDipoleForceAndEnergy, :3871-4024
    double recip_induced(std::vector<V3>& F, std::vector<V3>& T) const {
        double energy = 0;
        int ce[20][3];
        for (int k = 0; k < 20; k++) comp_exp(k, ce[k]);
        bool mutual = S.polarization == 0;
        for (int i = 0; i < n; i++) {
            const Site& s = site[i];
            double c0, c2[6], c3[10];
            V3 c1;
            cartesian_potential(phidp, i, c0, c1, c2, c3);
            V3 d = s.d;
            if (mutual && s.aniso) d += mu[i];
            T[i] += multipole_torque(d, s.Q, s.O, c1, c2, c3) * kElectric;
            double uf[3];
            for (int j = 0; j < 3; j++) uf[j] = A[0][j] * mu[i][0] + A[1][j] * mu[i][1] + A[2][j] * mu[i][2];
            const double* ph = &phi[35 * (size_t)i];
            const double* pd = &phidp[35 * (size_t)i];
            for (int k = 0; k < 3; k++) energy += 2.0 * uf[k] * ph[1 + k];
            V3 f;
            for (int k = 0; k < 3; k++)
                for (int a = 0; a < 3; a++) {
                    int e2[3] = {0, 0, 0};
                    e2[k]++; e2[a]++;
                    int id = phi_index(e2[0], e2[1], e2[2]);
                    f[a] += 2.0 * uf[k] * ph[id];
                    if (mutual) f[a] += 2.0 * uf[k] * pd[id];
                }
            const double* m = &mfrac[20 * (size_t)i];
            for (int k = 0; k < 20; k++)
                for (int a = 0; a < 3; a++) {
                    int e2[3] = {ce[k][0], ce[k][1], ce[k][2]};
                    e2[a]++;
                    f[a] += 2.0 * m[k] * pd[phi_index(e2[0], e2[1], e2[2])];
                }
            f = f * (0.5 * kElectric);
            for (int a = 0; a < 3; a++) F[i][a] -= A[a][0] * f[0] + A[a][1] * f[1] + A[a][2] * f[2];
        }
        return 0.25 * kElectric * energy;
    }

    // calculatePmeSelfEnergy :4283-4318, calculatePmeSelfTorque :4320-4333
    double self_energy() const {
        double cii = 0, dii = 0, qii = 0, oii = 0;
        for (int i = 0; i < n; i++) {
            const Site& s = site[i];
            cii += s.q * s.q;
            V3 d(s.sd[1], s.sd[2], s.sd[0]);
            dii += d.dot(d + mu[i]);
            for (int k = 0; k < 5; k++) qii += s.sq[k] * s.sq[k];
            for (int k = 0; k < 7; k++) oii += s.so[k] * s.so[k];
        }
        double pre = -S.alpha * kElectric / kSqrtPi;
        double a2 = S.alpha * S.alpha, a4 = a2 * a2, a6 = a4 * a2;
        return pre * (cii + (2.0 / 3.0) * a2 * dii + (4.0 / 15.0) * a4 * qii + (8.0 / 105.0) * a6 * oii);
    }
    void self_torque(std::vector<V3>& T) const {
        double term = (2.0 / 3.0) * kElectric * (S.alpha * S.alpha * S.alpha) / kSqrtPi;
        for (int i = 0; i < n; i++) {
            const Site& s = site[i];
            if (s.aniso) continue;
            V3 d(s.sd[1], s.sd[2], s.sd[0]);
            T[i] += d.cross(mu[i] * 2.0) * term;
        }
    }

    // ThreeFold frame by the exact chain rule (flag 1; the reference's own
    // mapping at :2068-2097 is an approximate projection).  The frame is
    // z = s/|s|, s = u + v + w (unit vectors to the z, x and y atoms), and
    // x = t/|t|, t = v - (v.z) z.  Moving an atom turns the frame by
    // dtheta = z x dz + (y.dx) z, so with tau = -dE/dtheta the force on atom a
    // is F_a = (dz/dr_a)^T (tau x z) + (tau.z) (dx/dr_a)^T y.
    void threefold_exact(int i, const V3& tq, std::vector<V3>& F) const {
        const V3 pI = site[i].pos;
        const int iz = S.atom_z[i], ix = S.atom_x[i], iy = S.atom_y[i];
        V3 u = site[iz].pos - pI, v = site[ix].pos - pI, w = site[iy].pos - pI;
        const double nu = normalize(u), nv = normalize(v), nw = normalize(w);
        V3 z = u + v + w;
        const double ns = normalize(z);
        V3 x = v - z * v.dot(z);
        const double nt = normalize(x);
        const V3 y = z.cross(x);
        const V3 c = y * (tq.dot(z) / nt);             // through x: on t, hence on v and on z
        const V3 bz = tq.cross(z) - c * v.dot(z);       // everything that moves with z
        const V3 gs = (bz - z * z.dot(bz)) * (1.0 / ns);  // on s = u + v + w
        auto perp = [](const V3& e, const V3& g) { return g - e * e.dot(g); };
        const V3 FZ = perp(u, gs) * (1.0 / nu);
        const V3 FX = perp(v, gs + c) * (1.0 / nv);
        const V3 FY = perp(w, gs) * (1.0 / nw);
        F[iz] += FZ;
        F[ix] += FX;
        F[iy] += FY;
        F[i] -= FZ + FX + FY;
    }

    // mapTorqueToForceForParticle, :1895-2110
    void torque_to_force(int i, const V3& tq, std::vector<V3>& F) const {
        int t = S.axis_types[i];
        if (t == 5 || S.atom_z[i] < 0) return;
        if (t == 3 && (S.flags & 1) && S.atom_y[i] >= 0) {
            threefold_exact(i, tq, F);
            return;
        }
        const V3 pI = site[i].pos;
        int iz = S.atom_z[i], ix = S.atom_x[i], iy = S.atom_y[i];
        double nU, nV, nW;
        V3 U = site[iz].pos - pI;
        nU = normalize(U);
        V3 Vv;
        if (t == 4) {
            // ZOnly: second vector chosen as in the frame construction
            Vv = std::fabs(U[0]) < 0.866 ? V3(1, 0, 0) : V3(0, 1, 0);
        } else {
            Vv = site[ix].pos - pI;
        }
        nV = normalize(Vv);
        V3 W;
        if ((t == 2 || t == 3) && iy >= 0) W = site[iy].pos - pI; else W = U.cross(Vv);
        nW = normalize(W);
        V3 UV = Vv.cross(U), UW = W.cross(U), VW = W.cross(Vv);
        normalize(UV); normalize(UW); normalize(VW);
        double cUV = U.dot(Vv), sUV = std::sqrt(1.0 - cUV * cUV);
        double cUW = U.dot(W), sUW = std::sqrt(1.0 - cUW * cUW);
        double cVW = Vv.dot(W), sVW = std::sqrt(1.0 - cVW * cVW);
        double dU = -U.dot(tq), dV = -Vv.dot(tq), dW = -W.dot(tq);
        if (t == 0 || t == 1) {
            double f1 = dV / (nU * sUV), f2 = dW / nU, f3 = -dU / (nV * sUV), f4 = 0;
            if (t == 1) { f2 *= 0.5; f4 = 0.5 * dW / nV; }
            V3 fu = UV * f1 + UW * f2, fv = UV * f3 + VW * f4;
            F[iz] -= fu;
            F[ix] -= fv;
            F[i] += fu + fv;
        } else if (t == 2) {
            V3 R = Vv + W;
            V3 Sv = U.cross(R);
            double nR = normalize(R), nS = normalize(Sv);
            (void)nR; (void)nS;
            V3 UR = R.cross(U), US = Sv.cross(U), VS = Sv.cross(Vv), WS = Sv.cross(W);
            normalize(UR); normalize(US); normalize(VS); normalize(WS);
            double cUR = U.dot(R), sUR = std::sqrt(1.0 - cUR * cUR);
            double cVS = Vv.dot(Sv), sVS = std::sqrt(1.0 - cVS * cVS);
            double cWS = W.dot(Sv), sWS = std::sqrt(1.0 - cWS * cWS);
            V3 t1 = Vv - Sv * cVS, t2 = W - Sv * cWS;
            normalize(t1); normalize(t2);
            double ut1c = U.dot(t1), ut1s = std::sqrt(1.0 - ut1c * ut1c);
            double ut2c = U.dot(t2), ut2s = std::sqrt(1.0 - ut2c * ut2c);
            double dR = -R.dot(tq), dS = -Sv.dot(tq);
            double f1 = dR / (nU * sUR), f2 = dS / nU, f3 = dU / (nV * (ut1s + ut2s)), f4 = dU / (nW * (ut1s + ut2s));
            V3 fu = UR * f1 + US * f2;
            V3 fv = (Sv * sVS - t1 * cVS) * f3;
            V3 fw = (Sv * sWS - t2 * cWS) * f4;
            F[iz] -= fu;
            F[ix] -= fv;
            F[iy] -= fw;
            F[i] += fu + fv + fw;
        } else if (t == 3) {
            V3 du = (UW * (dW / (nU * sUW)) + UV * (dV / (nU * sUV)) - UW * (dU / (nU * sUW)) - UV * (dU / (nU * sUV))) * (1.0 / 3.0);
            V3 dv = (VW * (dW / (nV * sVW)) - UV * (dU / (nV * sUV)) - VW * (dV / (nV * sVW)) + UV * (dV / (nV * sUV))) * (1.0 / 3.0);
            V3 dw = (UW * (-dU / (nW * sUW)) - VW * (dV / (nW * sVW)) + UW * (dW / (nW * sUW)) + VW * (dW / (nW * sVW))) * (1.0 / 3.0);
            F[iz] -= du;
            F[ix] -= dv;
            if (iy >= 0) F[iy] -= dw;
            F[i] += du + dv + dw;
        } else if (t == 4) {
            V3 du = UV * (dV / (nU * sUV)) + UW * (dW / nU);
            F[iz] -= du;
            F[i] += du;
        }
    }
};

}  // namespace

extern "C" {

void oracle_bspline(double w, double* theta) { bspline(w, reinterpret_cast<double(*)[5]>(theta)); }

void oracle_bspline_moduli(int n, double* moduli) { bspline_moduli(n, moduli); }

int oracle_fft3d(double* data, int nx, int ny, int nz, int sign) {
    size_t tot = (size_t)nx * ny * nz;
    std::vector<cplx> g(tot);
    for (size_t i = 0; i < tot; i++) g[i] = cplx(data[2 * i], data[2 * i + 1]);
    fft3d(g, nx, ny, nz, sign);
    for (size_t i = 0; i < tot; i++) { data[2 * i] = g[i].real(); data[2 * i + 1] = g[i].imag(); }
    return 0;
}

// mapTorqueToForce (:2112-2131) alone: the forces that the frame chain rule
// gives for the given per-site torques at the given positions (overwritten).
int oracle_torque_to_force(const oracle_system* sys, const double* pos, const double* torques, double* forces) {
    try {
        Oracle o(*sys);
        const int n = sys->n;
        o.site.resize(n);
        for (int i = 0; i < n; i++) o.site[i].pos = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        std::vector<V3> F(n);
        for (int i = 0; i < n; i++) o.torque_to_force(i, V3(torques[3 * i], torques[3 * i + 1], torques[3 * i + 2]), F);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) forces[3 * i + k] = F[i][k];
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

static void setup(Oracle& o, const double* pos) {
    o.validate();
    o.set_box();
    if (o.periodic)
        for (int k = 0; k < 3; k++) {
            o.moduli[k].resize(o.gn[k]);
            bspline_moduli(o.gn[k], o.moduli[k].data());
        }
    o.load_and_rotate(pos);
    o.build_pairs();
    if (o.periodic) o.compute_splines();
}

int oracle_induced_field(const oracle_system* sys, const double* pos, const double* mu, double* field_out) {
    try {
        Oracle o(*sys);
        setup(o, pos);
        std::vector<V3> m(sys->n), E;
        for (int i = 0; i < sys->n; i++) m[i] = V3(mu[3 * i], mu[3 * i + 1], mu[3 * i + 2]);
        o.induced_field(m, E);
        for (int i = 0; i < sys->n; i++)
            for (int k = 0; k < 3; k++) field_out[3 * i + k] = E[i][k];
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// calculateElectrostaticPotentialForParticleGridPoint and
// calculateElectrostaticPotential, MPIDReferenceForce.cpp:2464-2531.
int oracle_potential(const oracle_system* sys, const double* pos, int npts, const double* points, double* out) {
    try {
        Oracle o(*sys);
        setup(o, pos);
        o.fixed_field();
        o.solve();
        if (!o.converged) return 2;
        for (int g = 0; g < npts; g++) {
            const V3 pt(points[3 * g], points[3 * g + 1], points[3 * g + 2]);
            double phi = 0.0;
            for (int i = 0; i < o.n; i++) {
                const Site& s = o.site[i];
                V3 d = s.pos - pt;  // particle minus point, as the reference
                o.min_image(d);
                const double r2 = d.dot(d), r = std::sqrt(r2);
                const double rr1 = 1.0 / r, rr3 = rr1 * rr1 * rr1, rr5 = 3.0 * rr3 * rr1 * rr1;
                double qrr = 0.0;
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++) qrr += d[a] * s.Q[idx2(a, c)] * d[c];
                phi += s.q * rr1 - (s.d + o.mu[i]).dot(d) * rr3 + qrr * rr5;
            }
            out[g] = kElectric * phi;
        }
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// calculateMPIDSystemMultipoleMoments, MPIDReferenceForce.cpp:2352-2462.
int oracle_multipole_moments(const oracle_system* sys, const double* pos, const double* masses, double* out) {
    try {
        Oracle o(*sys);
        setup(o, pos);
        o.fixed_field();
        o.solve();
        if (!o.converged) return 2;
        double mt = 0.0;
        V3 com;
        for (int i = 0; i < o.n; i++) {
            mt += masses[i];
            com += o.site[i].pos * masses[i];
        }
        if (mt > 0.0) com = com * (1.0 / mt);
        double q = 0.0, t[6] = {0, 0, 0, 0, 0, 0};  // traced second moment xx xy xz yy yz zz
        V3 dpl;
        for (int i = 0; i < o.n; i++) {
            const Site& s = o.site[i];
            const V3 x = s.pos - com, d = s.d + o.mu[i];
            q += s.q;
            dpl += x * s.q + d;
            for (int a = 0; a < 3; a++)
                for (int c = a; c < 3; c++) t[idx2(a, c)] += x[a] * x[c] * s.q + x[a] * d[c] + x[c] * d[a];
        }
        const double tr = (t[idx2(0, 0)] + t[idx2(1, 1)] + t[idx2(2, 2)]) / 3.0;
        double Q[3][3];
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
                Q[a][c] = 0.5 * (t[idx2(a, c)] - (a == c ? tr : 0.0));
                for (int i = 0; i < o.n; i++) Q[a][c] += o.site[i].Q[idx2(a, c)];
            }
        const double debye = 4.80321;  // the reference's value (:2447)
        out[0] = q;
        for (int a = 0; a < 3; a++) out[1 + a] = dpl[a] * 10.0 * debye;
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) out[4 + 3 * a + c] = Q[a][c] * 100.0 * 3.0 * debye;
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

int oracle_evaluate(const oracle_system* sys, const double* pos, oracle_out* out) {
    try {
        Oracle o(*sys);
        int n = sys->n;
        setup(o, pos);
        o.fixed_field();
        o.solve();
        if (out) {
            out->iterations = o.iterations;
            out->eps = o.last_eps;
            out->converged = o.converged ? 1 : 0;
        }
        if (!o.converged) {
            if (out)
                std::snprintf(out->error, sizeof(out->error), "Induced dipoles did not converge:  iterations=%d eps=%g",
                              o.iterations, o.last_eps);
            return 2;
        }
        // final induced reciprocal potential (phidp) from the converged dipoles
        std::vector<V3> Ei;
        o.induced_field(o.mu, Ei, &o.phidp);
        std::vector<V3> F(n), T(n);
        double ereal = 0;
        for (const auto& p : o.pairs) ereal += o.electro_pair(p, F, T);
        o.self_torque(T);
        double eind = o.recip_induced(F, T);
        double efix = o.recip_fixed(F, T);
        double eself = o.self_energy();
        if (sys->polarization == 2) o.extrapolation_response(F, T);
        if (out) {
            std::vector<V3> T0 = T;
            for (int i = 0; i < n; i++) o.torque_to_force(i, T[i], F);
            out->energy_terms[0] = ereal;
            out->energy_terms[1] = efix;
            out->energy_terms[2] = eind;
            out->energy_terms[3] = eself;
            out->energy = ereal + efix + eind + eself;
            out->energy_variational = out->energy;
            if (sys->polarization == 0) {
                // residual r = E0 + T mu - alpha^-1 mu of the solution (mu = 0 where alpha = 0)
                double mr = 0.0;
                for (int i = 0; i < n; i++) {
                    const double* a = o.site[i].labpol;
                    const double A[3][3] = {{a[0], a[1], a[2]}, {a[1], a[3], a[4]}, {a[2], a[4], a[5]}};
                    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
                    if (det == 0.0) continue;
                    // alpha^-1 mu by Cramer's rule
                    const V3& m = o.mu[i];
                    double x[3];
                    for (int c = 0; c < 3; c++) {
                        double B[3][3];
                        for (int r = 0; r < 3; r++)
                            for (int k = 0; k < 3; k++) B[r][k] = (k == c) ? m[r] : A[r][k];
                        x[c] = (B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                                B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                                B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0])) / det;
                    }
                    for (int k = 0; k < 3; k++) mr += m[k] * (o.E0[i][k] + Ei[i][k] - x[k]);
                }
                out->energy_variational -= 0.5 * kElectric * mr;
            }
            for (int i = 0; i < n; i++) {
                const Site& s = o.site[i];
                for (int k = 0; k < 3; k++) {
                    if (out->lab_dipoles) out->lab_dipoles[3 * i + k] = s.d[k];
                    if (out->fixed_field) out->fixed_field[3 * i + k] = o.E0[i][k];
                    if (out->induced) out->induced[3 * i + k] = o.mu[i][k];
                    if (out->torques) out->torques[3 * i + k] = T0[i][k];
                    if (out->forces) out->forces[3 * i + k] = F[i][k];
                }
                for (int k = 0; k < 6; k++) {
                    if (out->lab_quads) out->lab_quads[6 * i + k] = s.Q[k];
                    if (out->lab_pol) out->lab_pol[6 * i + k] = s.labpol[k];
                }
                for (int k = 0; k < 10; k++)
                    if (out->lab_octs) out->lab_octs[10 * i + k] = s.O[k];
                for (int k = 0; k < 35; k++) {
                    if (out->phi) out->phi[35 * i + k] = o.phi[35 * i + k];
                    if (out->phidp) out->phidp[35 * i + k] = o.phidp[35 * i + k];
                }
            }
        }
        return 0;
    } catch (const std::exception& e) {
        if (out) std::snprintf(out->error, sizeof(out->error), "%s", e.what());
        return 1;
    }
}

}  // extern "C"
