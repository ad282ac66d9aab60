// Host check of the pair formulas the energy gradients use (pair_math.h,
// instantiated here in double precision), against what linear algebra and
// calculus fix:
//
//  qi_transpose    qi_potential_to_lab is the transpose of qi_moments:
//                  <V, L m> = <L^T V, m> for random frames, moments and V.
//  pot_general     the potential vectors of elec_pair<kPot> are the
//                  derivatives, at fixed induced dipoles, of the pair energy
//                  in which the induced dipoles interact fully with the
//                  moments.  That energy is elec_pair's with the dipoles
//                  doubled (its induced terms carry one half); it is linear
//                  in each moment, so central differences are exact to
//                  rounding.
//  pot_bare        the same for elec_pair_bare (a's components and b's charge).
//  thole           thole_u_derivative against central differences, in the
//                  Thole width a, of the damped energy assembled from the
//                  field formulas: -mu_i.E_i - mu_j.E_j - mu_i.T.mu_j.
//
// Prints the largest relative difference of each check and its magnitude.
// Build and run:  make -C mpidopenmmplugin_amd/csrc pair_adjoint
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>

#include "pair_math.h"

using namespace mpid;

static PairParams<double> params(int fix_torque) {
    PairParams<double> P{};
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
    P.default_thole = 1.3;
    P.scale14 = 1.0;
    P.mutual = 1;
    P.fix_torque = fix_torque;
    return P;
}

namespace {
struct MaxRel {
    double diff = 0.0, mag = 0.0;
    void add(double a, double b) {
        diff = std::max(diff, std::fabs(a - b));
        mag = std::max(mag, std::max(std::fabs(a), std::fabs(b)));
    }
    double rel() const { return mag > 0.0 ? diff / mag : diff; }
};
}  // namespace

int main() {
    std::mt19937_64 rng(20261203);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    MaxRel qt, pg, pb, th;
    auto random_moments = [&](double* m) {
        m[0] = 0.8 * U(rng);
        for (int k = 1; k < 4; k++) m[k] = 0.02 * U(rng);
        for (int k = 4; k < 10; k++) m[k] = 1e-3 * U(rng);
        for (int k = 10; k < kMP; k++) m[k] = 1e-4 * U(rng);
    };
    for (int trial = 0; trial < 2000; trial++) {
        const PairParams<double> P = params(trial & 1);
        V3<double> dr;
        const double len = 0.12 + 0.7 * (0.5 + 0.5 * U(rng));
        if (trial % 97 == 0) {
            dr = mk(len, 0.0, 0.0);
        } else {
            dr = mk(U(rng), U(rng), U(rng));
            dr = dr * (len / std::sqrt(dot(dr, dr)));
        }
        const double r2 = dot(dr, dr);
        double ma[kMP], mb[kMP];
        random_moments(ma);
        random_moments(mb);

        // transpose identity
        {
            V3<double> ex, ey, ez;
            qi_frame(dr, 1.0 / std::sqrt(r2), ex, ey, ez);
            double q[16], V[16], g[kMP] = {};
            qi_moments(ma, ex, ey, ez, q);
            for (int k = 0; k < 16; k++) V[k] = U(rng);
            qi_potential_to_lab(V, ex, ey, ez, g);
            double lhs = 0.0, rhs = 0.0, scale = 0.0;
            for (int k = 0; k < 16; k++) {
                lhs += V[k] * q[k];
                scale += std::fabs(V[k] * q[k]);
            }
            for (int k = 0; k < kMP; k++) rhs += g[k] * ma[k];
            qt.diff = std::max(qt.diff, std::fabs(lhs - rhs) / scale);
            qt.mag = std::max(qt.mag, scale);
        }

        const V3<double> mua = mk(0.01 * U(rng), 0.01 * U(rng), 0.01 * U(rng));
        const V3<double> mub = mk(0.01 * U(rng), 0.01 * U(rng), 0.01 * U(rng));
        const double dmpa = 0.3 + 0.05 * U(rng), dmpb = (trial % 3 == 0) ? 0.0 : 0.2 + 0.05 * U(rng);
        const R2<double> ta{0.39, dmpa}, tb{0.41, dmpb};
        const bool anisoA = (trial % 4) >= 2, anisoB = (trial % 8) >= 4;
        const double ms = (trial % 5 == 0) ? 0.5 : 1.0;
        const double ps = (trial % 7 == 0) ? 0.0 : ((trial % 7 == 1) ? 0.5 : 1.0);
        const V3<double> mua2 = mua * 2.0, mub2 = mub * 2.0;

        // general pair: potentials against central differences of the energy
        {
            double pa[kMP] = {}, pbv[kMP] = {};
            elec_pair<double, true>(dr, r2, ma, mb, mua, mub, ta, tb, anisoA, anisoB, ms, ps, P, true, pa, pbv);
            for (int side = 0; side < 2; side++)
                for (int k = 0; k < kMP; k++) {
                    double a1[kMP], b1[kMP], a2[kMP], b2[kMP];
                    std::copy(ma, ma + kMP, a1);
                    std::copy(mb, mb + kMP, b1);
                    std::copy(ma, ma + kMP, a2);
                    std::copy(mb, mb + kMP, b2);
                    const double h = 1e-3;
                    (side == 0 ? a1 : b1)[k] += h;
                    (side == 0 ? a2 : b2)[k] -= h;
                    const double ep = elec_pair<double>(dr, r2, a1, b1, mua2, mub2, ta, tb, anisoA, anisoB, ms, ps, P)
                                          .energy;
                    const double em = elec_pair<double>(dr, r2, a2, b2, mua2, mub2, ta, tb, anisoA, anisoB, ms, ps, P)
                                          .energy;
                    pg.add((ep - em) / (2.0 * h), (side == 0 ? pa : pbv)[k]);
                }
        }
        // one bare charge b: a's potential (no charge-charge channel) and b's
        {
            double pa[kMP] = {}, pb0 = 0.0;
            const double qb = mb[0];
            elec_pair_bare<double, true>(dr, r2, ma, qb, mua, ta, tb, anisoA, ms, ps, P, pa, &pb0);
            for (int k = 0; k < kMP; k++) {
                double a1[kMP], a2[kMP];
                std::copy(ma, ma + kMP, a1);
                std::copy(ma, ma + kMP, a2);
                const double h = 1e-3;
                a1[k] += h;
                a2[k] -= h;
                const double ep = elec_pair_bare<double>(dr, r2, a1, qb, mua2, ta, tb, anisoA, ms, ps, P).energy;
                const double em = elec_pair_bare<double>(dr, r2, a2, qb, mua2, ta, tb, anisoA, ms, ps, P).energy;
                pb.add((ep - em) / (2.0 * h), pa[k]);
            }
            const double h = 1e-3;
            const double ep = elec_pair_bare<double>(dr, r2, ma, qb + h, mua2, ta, tb, anisoA, ms, ps, P).energy;
            const double em = elec_pair_bare<double>(dr, r2, ma, qb - h, mua2, ta, tb, anisoA, ms, ps, P).energy;
            pb.add((ep - em) / (2.0 * h), pb0);
        }
        // Thole width: both sites polarisable (the derivative vanishes otherwise)
        {
            const double da = 0.28 + 0.05 * U(rng), db = 0.25 + 0.05 * U(rng);
            const double a = 0.8 + 0.5 * (0.5 + 0.5 * U(rng));
            const double ds = (trial % 3 == 0) ? 0.0 : ((trial % 3 == 1) ? 0.5 : 1.0);
            const bool mutual = (trial % 2) == 0;
            auto energy = [&](double aw) {
                // fixed_field_radial takes the width from the sites' Thole
                // factors when ds = 0, else from the default width
                PairParams<double> Q = P;
                Q.default_thole = aw;
                const R2<double> ti{0.5 * aw, da}, tj{0.5 * aw, db};
                V3<double> fi, fj;
                fixed_field_pair<double>(dr, r2, ma, mb, ti, tj, ds, Q, fi, fj);
                double e = -(dot(mua, fi) + dot(mub, fj));
                if (mutual) e -= dot(mua, induced_field_from<double>(mub, dr, induced_field_radial<double>(r2, da, db, aw, Q)));
                return kElectric * e;
            };
            const double hstep = 1e-5 * a;
            const double fd = a * (energy(a + hstep) - energy(a - hstep)) / (2.0 * hstep);
            const double an = thole_u_derivative<double>(dr, r2, ma, mb, mua, mub, da, db, a, ds, mutual);
            th.add(fd, an);
        }
    }
    std::printf("qi_transpose_rel %.3e\n", qt.diff);
    std::printf("pot_general_rel %.3e\npot_general_mag %.3e\n", pg.rel(), pg.mag);
    std::printf("pot_bare_rel %.3e\npot_bare_mag %.3e\n", pb.rel(), pb.mag);
    std::printf("thole_rel %.3e\nthole_mag %.3e\n", th.rel(), th.mag);
    return 0;
}
