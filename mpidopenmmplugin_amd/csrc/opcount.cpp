// Algorithmic floating-point operation counts of the pair formulas.
//
// The pair functions of pair_math.h are templates; here they are instantiated
// on the host with a number type that carries a double value and counts every
// arithmetic operation it takes part in: +, -, * and / count one each (so a
// fused multiply-add counts two), exp, erfc and reciprocal square root one
// each; sign changes, comparisons and selects count nothing.  Each function
// is evaluated for one pair in both Thole branches (both sites polarisable,
// so the damping terms are evaluated; or a damping factor of zero), which
// are the two costs the roofline weights by the pair mix.
//
// Prints one line per formula:  name  flops_damped  flops_undamped
// (the field formulas one-sided: radial factors of the pair, and the field
// that one site's moments produce at the other)
// Build and run:  make -C mpidopenmmplugin_amd/csrc opcount
#include <cmath>
#include <cstdio>

#include "pair_math.h"

namespace {

struct Ops {
    static long n;
};
long Ops::n = 0;

struct Cnt {
    double v = 0.0;
    Cnt() = default;
    Cnt(double x) : v(x) {}
    explicit operator double() const { return v; }
};
inline Cnt operator+(Cnt a, Cnt b) { Ops::n++; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { Ops::n++; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { Ops::n++; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { Ops::n++; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt& operator+=(Cnt& a, Cnt b) { a = a + b; return a; }
inline Cnt& operator-=(Cnt& a, Cnt b) { a = a - b; return a; }
inline Cnt& operator*=(Cnt& a, Cnt b) { a = a * b; return a; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator==(Cnt a, Cnt b) { return a.v == b.v; }
inline bool operator!=(Cnt a, Cnt b) { return a.v != b.v; }
inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }
inline Cnt fabs(Cnt a) { return Cnt(std::fabs(a.v)); }
inline Cnt fmin(Cnt a, Cnt b) { return Cnt(std::fmin(a.v, b.v)); }
inline Cnt floor(Cnt a) { Ops::n++; return Cnt(std::floor(a.v)); }
// found by argument-dependent lookup from the pair templates
Cnt rsqrt_r(Cnt x) { Ops::n++; return Cnt(1.0 / std::sqrt(x.v)); }
Cnt exp_r(Cnt x) { Ops::n++; return Cnt(std::exp(x.v)); }
Cnt erfc_r(Cnt x, Cnt) { Ops::n++; return Cnt(std::erfc(x.v)); }
Cnt div_r(Cnt a, Cnt b) { return a / b; }

}  // namespace

namespace mpid {
template <> struct R2T<Cnt> {
    struct type {
        Cnt x, y;
    };
};
}  // namespace mpid

using namespace mpid;

static PairParams<Cnt> params(bool mutual) {
    PairParams<Cnt> P{};
    const double alpha = 2.9203, rc = 0.9;
    P.box.L[0] = P.box.L[1] = P.box.L[2] = 9.932;
    P.box.inv[0] = P.box.inv[1] = P.box.inv[2] = 1.0 / 9.932;
    P.n = 2;
    P.rc2 = rc * rc;
    P.alpha = alpha;
    double aln = 1.0 / (kSqrtPi * alpha);
    P.aln[0] = 0.0;
    for (int k = 1; k <= 4; k++) {
        aln *= 2.0 * alpha * alpha;
        P.aln[k] = aln;
    }
    P.default_thole = 5.0;
    P.scale14 = 1.0;
    P.mutual = mutual ? 1 : 0;
    P.fix_torque = 0;
    return P;
}

int main() {
    // an intermolecular pair at 0.3 nm with generic moments
    const V3<Cnt> dr = mk(Cnt(0.17), Cnt(-0.21), Cnt(0.12));
    const Cnt r2 = dr.x.v * dr.x.v + dr.y.v * dr.y.v + dr.z.v * dr.z.v;
    Cnt ma[kMP], mb[kMP];
    for (int k = 0; k < kMP; k++) {
        ma[k] = 0.01 * (k + 1);
        mb[k] = -0.013 * (k + 2);
    }
    const V3<Cnt> ua = mk(Cnt(0.004), Cnt(-0.002), Cnt(0.003)), ub = mk(Cnt(-0.001), Cnt(0.005), Cnt(0.002));
    const PairParams<Cnt> P = params(true);
    // one-sided pieces (radial factors, field from one site) and the
    // two-sided electrostatics; radial factors with and without damping
    long frad[2], ffrom = 0, irad[2], ifrom = 0, elec[2], bare[2], cart[2];
    for (int d = 0; d < 2; d++) {
        const double damp = d == 0 ? 0.316 : 0.0;  // d = 0: both sites polarisable
        R2<Cnt> ta{Cnt(1.3), Cnt(damp)}, tb{Cnt(1.3), Cnt(0.316)};
        Ops::n = 0;
        const FixedRadial<Cnt> g = fixed_field_radial<Cnt>(r2, ta, tb, Cnt(1.0), P);
        frad[d] = Ops::n;
        Ops::n = 0;
        (void)fixed_field_from<Cnt>(mb, dr, g);
        ffrom = Ops::n;
        Ops::n = 0;
        const InducedRadial<Cnt> f = induced_field_radial<Cnt>(r2, ta.y, tb.y, P.default_thole, P);
        irad[d] = Ops::n;
        Ops::n = 0;
        (void)induced_field_from<Cnt>(ub, dr, f);
        ifrom = Ops::n;
        Ops::n = 0;
        // the charge-charge channel is counted apart (charge_pair)
        ElecOut<Cnt> o = elec_pair<Cnt>(dr, r2, ma, mb, ua, ub, ta, tb, true, true, Cnt(1.0), Cnt(1.0), P, false);
        (void)o;
        elec[d] = Ops::n;
        // a pair with one bare charge (b: charge only, no polarisability)
        Ops::n = 0;
        ElecBareOut<Cnt> ob = elec_pair_bare<Cnt>(dr, r2, ma, mb[0], ua, ta, tb, true, Cnt(1.0), Cnt(1.0), P);
        (void)ob;
        bare[d] = Ops::n;
        // its Cartesian form at scale factors 1, which the tile kernel runs
        Ops::n = 0;
        ElecBareOut<Cnt> oc = elec_pair_bare_cart<Cnt>(dr, r2, ma, mb[0], ua, true, P);
        (void)oc;
        cart[d] = Ops::n;
    }
    V3<Cnt> fc;
    Cnt ec;
    Ops::n = 0;
    charge_pair<Cnt>(dr, r2, Cnt(0.4), Cnt(0.4), P, fc, ec);
    const long cc = Ops::n;
    std::printf("fixed_field_radial %ld %ld\n", frad[0], frad[1]);
    std::printf("fixed_field_from %ld %ld\n", ffrom, ffrom);
    std::printf("induced_field_radial %ld %ld\n", irad[0], irad[1]);
    std::printf("induced_field_from %ld %ld\n", ifrom, ifrom);
    std::printf("elec_pair %ld %ld\n", elec[0], elec[1]);
    std::printf("elec_pair_bare %ld %ld\n", bare[0], bare[1]);
    std::printf("elec_pair_bare_cart %ld %ld\n", cart[0], cart[1]);
    std::printf("charge_pair %ld %ld\n", cc, cc);
    return 0;
}
