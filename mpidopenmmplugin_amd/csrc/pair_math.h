// Pair interaction formulas shared by the tile, neighbour-list and covalent
// pair kernels (device functions, inlined into each kernel).
//
// Physics follows MPIDReferencePmeForce
// (platforms/reference/src/SimTKReference/MPIDReferenceForce.cpp):
//   fixed field     :2812-2920 (+ getDampedInverseDistances :2678-2718)
//   induced field   :4161-4281
//   electrostatics  :4335-4920 (quasi-internal frame)
#pragma once

#include "engine.h"

namespace mpid {

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------
// Scale of a covalent partner entry (partner << 2 | class).
template <class R> MPID_HD R partner_scale(int e, R s14) {
    return ((e & 3) == kScaleZero) ? R(0) : s14;
}

// Quadrupole times vector and octopole contracted twice with a vector.
template <class R> MPID_HD V3<R> quad_dot(const R* Q, V3<R> v) {
    return {Q[0] * v.x + Q[1] * v.y + Q[2] * v.z, Q[1] * v.x + Q[3] * v.y + Q[4] * v.z,
            Q[2] * v.x + Q[4] * v.y + Q[5] * v.z};
}
// M = O . v (symmetric 3x3, stored XX XY XZ YY YZ ZZ)
template <class R> MPID_HD void oct_dot(const R* O, V3<R> v, R* M) {
    M[0] = O[0] * v.x + O[1] * v.y + O[2] * v.z;  // xx.
    M[1] = O[1] * v.x + O[3] * v.y + O[4] * v.z;  // xy.
    M[2] = O[2] * v.x + O[4] * v.y + O[5] * v.z;  // xz.
    M[3] = O[3] * v.x + O[6] * v.y + O[7] * v.z;  // yy.
    M[4] = O[4] * v.x + O[7] * v.y + O[8] * v.z;  // yz.
    M[5] = O[5] * v.x + O[8] * v.y + O[9] * v.z;  // zz.
}
template <class R> MPID_HD R quad_form(const R* M, V3<R> a, V3<R> b) {
    return dot(a, quad_dot(M, b));
}

// ---------------------------------------------------------------------------
// fixed multipole field
// ---------------------------------------------------------------------------
// Radial factors of the permanent field of a pair at distance r^2 = r2: the
// Ewald terms bn1..bn4 minus the Thole-damped bare terms (1 - ds lambda_k)
// (2k-1)!! / r^(2k+1), ds the multipole/polarization scale of the pair
// (MPIDReferenceForce.cpp:2812-2920, getDampedInverseDistances :2678-2718).
template <class R> struct FixedRadial {
    R g1, g2, g3, g4;
};
template <class R>
MPID_HD FixedRadial<R> fixed_field_radial(R r2, R2<R> tdi, R2<R> tdj, R ds, const PairParams<R>& P) {
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    const R rinv2 = rinv * rinv;
    const R ar = P.alpha * r;
    const R ex2 = exp_r(-ar * ar);
    const R bn0 = erfc_r(ar, ex2) * rinv;
    const R bn1 = (bn0 + P.aln[1] * ex2) * rinv2;
    const R bn2 = (R(3) * bn1 + P.aln[2] * ex2) * rinv2;
    const R bn3 = (R(5) * bn2 + P.aln[3] * ex2) * rinv2;
    const R bn4 = (R(7) * bn3 + P.aln[4] * ex2) * rinv2;
    R l3 = R(1), l5 = R(1), l7 = R(1), l9 = R(1);
    const R damp = tdi.y * tdj.y;
    if (damp != R(0)) {
        const R a = (ds == R(0)) ? tdi.x + tdj.x : P.default_thole;
        const R u = div_r(a * r, damp);
        if (u < R(50)) {
            const R e = exp_r(-u), u2 = u * u, u3 = u2 * u, u4 = u3 * u, u5 = u4 * u;
            const R p3 = R(1) + u + R(0.5) * u2;
            const R p5 = p3 + u3 * R(1.0 / 6.0);
            const R p7 = p5 + u4 * R(1.0 / 30.0);
            const R p9 = p5 + u4 * R(4.0 / 105.0) + u5 * R(1.0 / 210.0);
            l3 = R(1) - e * p3;
            l5 = R(1) - e * p5;
            l7 = R(1) - e * p7;
            l9 = R(1) - e * p9;
        }
    }
    const R rinv3 = rinv * rinv2, rinv5 = rinv3 * rinv2, rinv7 = rinv5 * rinv2, rinv9 = rinv7 * rinv2;
    FixedRadial<R> g;
    g.g1 = bn1 - (R(1) - ds * l3) * rinv3;
    g.g2 = bn2 - R(3) * (R(1) - ds * l5) * rinv5;
    g.g3 = bn3 - R(15) * (R(1) - ds * l7) * rinv7;
    g.g4 = bn4 - R(105) * (R(1) - ds * l9) * rinv9;
    return g;
}

// Field at site i from the permanent moments m of a site at r_i + dr.  The
// field at that site from the moments of i is fixed_field_from(m_i, -dr, g).
template <class R> MPID_HD V3<R> fixed_field_from(const R* m, V3<R> dr, const FixedRadial<R>& g) {
    const V3<R> d = mk(m[1], m[2], m[3]);
    const V3<R> q = quad_dot(m + 4, dr);
    R M[6];
    oct_dot(m + 10, dr, M);
    const V3<R> o = quad_dot(M, dr);
    const R dr_d = dot(d, dr), dr_q = dot(q, dr), dr_o = dot(o, dr);
    return q * (R(2) * g.g2) - d * g.g1 - dr * (g.g1 * m[0] - g.g2 * dr_d + g.g3 * dr_q - g.g4 * dr_o) -
           o * (R(3) * g.g3);
}

// Both sides: field at i from j (fi) and at j from i (fj), dr = r_j - r_i.
template <class R>
MPID_HD void fixed_field_pair(V3<R> dr, R r2, const R* mi, const R* mj, R2<R> tdi, R2<R> tdj, R ds,
                              const PairParams<R>& P, V3<R>& fi, V3<R>& fj) {
    const FixedRadial<R> g = fixed_field_radial<R>(r2, tdi, tdj, ds, P);
    fi = fixed_field_from<R>(mj, dr, g);
    fj = fixed_field_from<R>(mi, mk(-dr.x, -dr.y, -dr.z), g);
}

// Lanes of block J that are covalent partners of the lane's atom: those pairs
// are left to the covalent-pair kernels.
__device__ __forceinline__ unsigned long long partner_mask(const int* __restrict__ parts, int np, int J) {
    unsigned long long m = 0;
    for (int k = 0; k < np; k++) {
        const int j = parts[k] >> 2;
        if ((j >> 6) == J) m |= 1ull << (j & 63);
    }
    return m;
}

// ---------------------------------------------------------------------------
// induced dipole field (the real-space part of the solver's matrix-vector product)
// ---------------------------------------------------------------------------
// MPIDReferenceForce.cpp:4161-4281.  Radial factors f1, f2 of the dipole
// field tensor of a pair: field at i from a dipole u at r_i + dr is
// dr (u.dr) f2 + u f1 (symmetric in the two sites).  dampi, dampj: damping
// factors; a: the Thole width of the pair (the sum of the sites' Thole
// factors for pairs with polarization scale 0, the default width otherwise).
template <class R> struct InducedRadial {
    R f1, f2;
};
template <class R>
MPID_HD InducedRadial<R> induced_field_radial(R r2, R dampi, R dampj, R a, const PairParams<R>& P) {
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    const R rinv2 = rinv * rinv;
    const R ar = P.alpha * r;
    const R ex2 = exp_r(-ar * ar);
    const R bn0 = erfc_r(ar, ex2) * rinv;
    const R bn1 = (bn0 + P.aln[1] * ex2) * rinv2;
    const R bn2 = (R(3) * bn1 + P.aln[2] * ex2) * rinv2;
    // (1 - lambda) of the Thole damping, computed directly
    R m3 = R(0), m5 = R(0);
    const R damp = dampi * dampj;
    if (damp != R(0)) {
        const R u = div_r(a * r, damp);
        if (u < R(50)) {
            const R e = exp_r(-u);
            m3 = e * (R(1) + u + R(0.5) * u * u);
            m5 = m3 + e * u * u * u * R(1.0 / 6.0);
        }
    }
    const R rinv3 = rinv * rinv2;
    InducedRadial<R> f;
    f.f1 = m3 * rinv3 - bn1;
    f.f2 = bn2 - R(3) * m5 * rinv3 * rinv2;
    return f;
}
template <class R> MPID_HD V3<R> induced_field_from(V3<R> u, V3<R> dr, const InducedRadial<R>& f) {
    return dr * (dot(u, dr) * f.f2) + u * f.f1;
}

// The field factors with the third one of the field gradient (Extrapolated
// polarization): f3 = bn3 - 15 (1 - lambda7) / r^7 with the Thole factor
// lambda7 of MPIDReferenceForce.cpp:4214 (the damped pair terms of
// calculateDirectInducedDipolePairIxns, :4161-4228).
template <class R> struct InducedRadialGrad {
    R f1, f2, f3;
};
template <class R>
MPID_HD InducedRadialGrad<R> induced_grad_radial(R r2, R dampi, R dampj, R a, const PairParams<R>& P) {
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    const R rinv2 = rinv * rinv;
    const R ar = P.alpha * r;
    const R ex2 = exp_r(-ar * ar);
    const R bn0 = erfc_r(ar, ex2) * rinv;
    const R bn1 = (bn0 + P.aln[1] * ex2) * rinv2;
    const R bn2 = (R(3) * bn1 + P.aln[2] * ex2) * rinv2;
    const R bn3 = (R(5) * bn2 + P.aln[3] * ex2) * rinv2;
    R m3 = R(0), m5 = R(0), m7 = R(0);
    const R damp = dampi * dampj;
    if (damp != R(0)) {
        const R u = div_r(a * r, damp);
        if (u < R(50)) {
            const R e = exp_r(-u);
            const R u2 = u * u;
            m3 = e * (R(1) + u + R(0.5) * u2);
            m5 = m3 + e * u2 * u * R(1.0 / 6.0);
            m7 = m5 + e * u2 * u2 * R(1.0 / 30.0);
        }
    }
    const R rinv3 = rinv * rinv2, rinv5 = rinv3 * rinv2;
    InducedRadialGrad<R> f;
    f.f1 = m3 * rinv3 - bn1;
    f.f2 = bn2 - R(3) * m5 * rinv5;
    f.f3 = bn3 - R(15) * m7 * rinv5 * rinv2;
    return f;
}
// Adds the gradient of the field of dipole u at separation dr = (source -
// field point) to g (xx yy zz xy xz yz), MPIDReferenceForce.cpp:4236-4248.
template <class R> MPID_HD void induced_grad_add(V3<R> u, V3<R> dr, R f2, R f3, R* g) {
    const R ud = dot(u, dr);
    const R t = ud * f3;
    g[0] += t * dr.x * dr.x - (R(2) * u.x * dr.x + ud) * f2;
    g[1] += t * dr.y * dr.y - (R(2) * u.y * dr.y + ud) * f2;
    g[2] += t * dr.z * dr.z - (R(2) * u.z * dr.z + ud) * f2;
    g[3] += t * dr.x * dr.y - (u.x * dr.y + u.y * dr.x) * f2;
    g[4] += t * dr.x * dr.z - (u.x * dr.z + u.z * dr.x) * f2;
    g[5] += t * dr.y * dr.z - (u.y * dr.z + u.z * dr.y) * f2;
}
template <class R>
MPID_HD void induced_field_pair(V3<R> dr, R r2, V3<R> ui, V3<R> uj, R dampi, R dampj, R a, const PairParams<R>& P,
                                V3<R>& fi, V3<R>& fj) {
    const InducedRadial<R> f = induced_field_radial<R>(r2, dampi, dampj, a, P);
    fi = induced_field_from<R>(uj, dr, f);
    fj = induced_field_from<R>(ui, dr, f);
}

// ---------------------------------------------------------------------------
// Energy gradient with respect to the Thole damping of a pair
// ---------------------------------------------------------------------------
// h = u dE/du for a pair of polarisable sites (kJ/mol), u = a r / (damp_i
// damp_j) the argument of the Thole damping and E the energy at fixed induced
// dipoles ui, uj.  The damping enters the permanent field that each site's
// induced dipole feels from the other (E contains -mu.E_perm, scale ds) and,
// with mutual polarization, the dipole-dipole coupling (E contains
// -1/2 mu.T mu, unscaled).  With lambda_k = 1 - exp(-u) P_k(u) as in
// fixed_field_radial: dlambda3/du = e u^2/2, dlambda5/du = e u^3/6,
// dlambda7/du = e (u^3 + u^4)/30, dlambda9/du = e (3u^3 + 3u^4 + u^5)/210.
// The chain rule then gives dE/da = h/a and dE/ddamp_i = -h/damp_i.
template <class R>
MPID_HD R thole_u_derivative(V3<R> dr, R r2, const R* mi, const R* mj, V3<R> ui, V3<R> uj, R dampi, R dampj,
                             R a, R ds, bool mutual) {
    const R damp = dampi * dampj;
    if (damp == R(0)) return R(0);
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    const R u = div_r(a * r, damp);
    if (!(u < R(50))) return R(0);
    const R e = exp_r(-u), u2 = u * u, u3 = u2 * u, u4 = u3 * u, u5 = u4 * u;
    const R dl3 = e * R(0.5) * u2, dl5 = e * u3 * R(1.0 / 6.0), dl7 = e * (u3 + u4) * R(1.0 / 30.0),
            dl9 = e * (R(3) * u3 + R(3) * u4 + u5) * R(1.0 / 210.0);
    const R rinv2 = rinv * rinv, rinv3 = rinv * rinv2, rinv5 = rinv3 * rinv2, rinv7 = rinv5 * rinv2,
            rinv9 = rinv7 * rinv2;
    R dE = R(0);
    if (ds != R(0)) {
        // d g_k/du of the permanent field's radial factors
        const FixedRadial<R> g{ds * dl3 * rinv3, R(3) * ds * dl5 * rinv5, R(15) * ds * dl7 * rinv7,
                               R(105) * ds * dl9 * rinv9};
        dE = dot(ui, fixed_field_from<R>(mj, dr, g)) + dot(uj, fixed_field_from<R>(mi, dr * R(-1), g));
    }
    if (mutual) {
        // d f1/du = -dlambda3/du / r^3, d f2/du = 3 dlambda5/du / r^5
        const InducedRadial<R> f{-dl3 * rinv3, R(3) * dl5 * rinv5};
        dE += dot(ui, induced_field_from<R>(uj, dr, f));
    }
    return -R(kElectric) * u * dE;
}

// ---------------------------------------------------------------------------
// Final real-space energy, force and torque in the quasi-internal (QI) frame.
// ---------------------------------------------------------------------------
// Spherical moments in a frame (ex, ey, ez): Cartesian tensors are contracted
// with the frame axes and projected as in MPIDReferenceForce.cpp:321-342.
template <class R>
MPID_HD void qi_moments(const R* m, V3<R> ex, V3<R> ey, V3<R> ez, R* q) {
    const R s3 = R(1.7320508075688772);
    q[0] = m[0];
    const V3<R> d = mk(m[1], m[2], m[3]);
    q[1] = dot(ez, d);
    q[2] = dot(ex, d);
    q[3] = dot(ey, d);
    const R* Q = m + 4;
    const V3<R> Qz = quad_dot(Q, ez), Qx = quad_dot(Q, ex), Qy = quad_dot(Q, ey);
    const R qzz = dot(ez, Qz), qxz = dot(ex, Qz), qyz = dot(ey, Qz);
    const R qxx = dot(ex, Qx), qyy = dot(ey, Qy), qxy = dot(ey, Qx);
    q[4] = R(3) * qzz;
    q[5] = R(2) * s3 * qxz;
    q[6] = R(2) * s3 * qyz;
    q[7] = s3 * (qxx - qyy);
    q[8] = R(2) * s3 * qxy;
    const R* O = m + 10;
    R Mz[6], Mx[6], My[6];
    oct_dot(O, ez, Mz);
    oct_dot(O, ex, Mx);
    oct_dot(O, ey, My);
    const V3<R> wz = quad_dot(Mz, ez);
    const R ozzz = dot(ez, wz), oxzz = dot(ex, wz), oyzz = dot(ey, wz);
    const R oxxz = quad_form(Mz, ex, ex), oyyz = quad_form(Mz, ey, ey), oxyz = quad_form(Mz, ex, ey);
    const R oxxx = quad_form(Mx, ex, ex), oxyy = quad_form(Mx, ey, ey), oxxy = quad_form(Mx, ex, ey);
    const R oyyy = quad_form(My, ey, ey);
    const R r32 = R(1.2247448713915890), r35 = R(0.7745966692414834), r110 = R(0.31622776601683794);
    q[9] = R(15) * ozzz;
    q[10] = R(15) * r32 * oxzz;
    q[11] = R(15) * r32 * oyzz;
    q[12] = R(15) * r35 * (oxxz - oyyz);
    q[13] = R(30) * r35 * oxyz;
    q[14] = R(15) * r110 * (oxxx - R(3) * oxyy);
    q[15] = R(15) * r110 * (R(3) * oxxy - oyyy);
}

// The transpose of qi_moments: given V_k = dE/dq_k for the 16 spherical QI
// moments, adds dE/dm_c for the 20 stored lab components m_c of the record
// (q, dipole, quadrupole XX XY XZ YY YZ ZZ, octopole XXX XXY XXZ XYY XYZ XZZ
// YYY YYZ YZZ ZZZ).  A stored off-diagonal component stands for every index
// ordering it represents, so its derivative sums the orderings.
template <class R> MPID_HD void sym2_add(R c, V3<R> a, V3<R> b, R* g) {
    // c (a_k b_l + a_l b_k)/2 summed over the orderings of each stored pair
    g[0] += c * a.x * b.x;
    g[1] += c * (a.x * b.y + a.y * b.x);
    g[2] += c * (a.x * b.z + a.z * b.x);
    g[3] += c * a.y * b.y;
    g[4] += c * (a.y * b.z + a.z * b.y);
    g[5] += c * a.z * b.z;
}
template <class R> MPID_HD void sym3_add(R c, V3<R> a, V3<R> b, V3<R> d, R* g) {
    // c a_k b_l d_m summed over the distinct orderings (k, l, m) of each stored triple
    const R ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z, dx = d.x, dy = d.y, dz = d.z;
    g[0] += c * ax * bx * dx;                                                       // xxx
    g[1] += c * (ax * bx * dy + ax * by * dx + ay * bx * dx);                       // xxy
    g[2] += c * (ax * bx * dz + ax * bz * dx + az * bx * dx);                       // xxz
    g[3] += c * (ax * by * dy + ay * bx * dy + ay * by * dx);                       // xyy
    g[4] += c * (ax * by * dz + ax * bz * dy + ay * bx * dz + ay * bz * dx + az * bx * dy + az * by * dx);  // xyz
    g[5] += c * (ax * bz * dz + az * bx * dz + az * bz * dx);                       // xzz
    g[6] += c * ay * by * dy;                                                       // yyy
    g[7] += c * (ay * by * dz + ay * bz * dy + az * by * dy);                       // yyz
    g[8] += c * (ay * bz * dz + az * by * dz + az * bz * dy);                       // yzz
    g[9] += c * az * bz * dz;                                                       // zzz
}
template <class R>
MPID_HD void qi_potential_to_lab(const R* V, V3<R> ex, V3<R> ey, V3<R> ez, R* g) {
    const R s3 = R(1.7320508075688772);
    g[0] += V[0];
    const V3<R> d = ez * V[1] + ex * V[2] + ey * V[3];
    g[1] += d.x;
    g[2] += d.y;
    g[3] += d.z;
    R* Q = g + 4;
    sym2_add(R(3) * V[4], ez, ez, Q);
    sym2_add(R(2) * s3 * V[5], ex, ez, Q);
    sym2_add(R(2) * s3 * V[6], ey, ez, Q);
    sym2_add(s3 * V[7], ex, ex, Q);
    sym2_add(-s3 * V[7], ey, ey, Q);
    sym2_add(R(2) * s3 * V[8], ey, ex, Q);
    const R r32 = R(1.2247448713915890), r35 = R(0.7745966692414834), r110 = R(0.31622776601683794);
    R* O = g + 10;
    sym3_add(R(15) * V[9], ez, ez, ez, O);
    sym3_add(R(15) * r32 * V[10], ex, ez, ez, O);
    sym3_add(R(15) * r32 * V[11], ey, ez, ez, O);
    sym3_add(R(15) * r35 * V[12], ez, ex, ex, O);
    sym3_add(-R(15) * r35 * V[12], ez, ey, ey, O);
    sym3_add(R(30) * r35 * V[13], ez, ex, ey, O);
    sym3_add(R(15) * r110 * V[14], ex, ex, ex, O);
    sym3_add(-R(45) * r110 * V[14], ex, ey, ey, O);
    sym3_add(R(45) * r110 * V[15], ex, ex, ey, O);
    sym3_add(-R(15) * r110 * V[15], ey, ey, ey, O);
}
// The same for a potential with only its m = 0 components (V1, V4, V9)
// nonzero besides the charge (a site facing a bare charge).
template <class R> MPID_HD void qi_potential_to_lab_m0(R V1, R V4, R V9, V3<R> ez, R* g) {
    g[1] += V1 * ez.x;
    g[2] += V1 * ez.y;
    g[3] += V1 * ez.z;
    sym2_add(R(3) * V4, ez, ez, g + 4);
    sym3_add(R(15) * V9, ez, ez, ez, g + 10);
}

// sum_i (dq_i/dtheta_k) V_i for rotations about the QI x, y and z axes.  The
// generator tables were derived symbolically from the Cartesian definition of
// the moments; they coincide with the hand-written tables at
// MPIDReferenceForce.cpp:4438-4455.
template <class R> MPID_HD void qi_torque_sums(const R* q, const R* V, R& tx, R& ty, R& tz) {
    const R s3 = R(1.7320508075688772), s6 = R(2.449489742783178), s52 = R(1.5811388300841898),
            s32 = R(1.2247448713915890);
    tx = q[3] * V[1] - q[1] * V[3] + s3 * q[6] * V[4] + q[8] * V[5] + (-s3 * q[4] - q[7]) * V[6] + q[6] * V[7] -
         q[5] * V[8] + s6 * q[11] * V[9] + s52 * q[13] * V[10] + (-s52 * q[12] - s6 * q[9]) * V[11] +
         (s52 * q[11] + s32 * q[15]) * V[12] + (-s52 * q[10] - s32 * q[14]) * V[13] + s32 * q[13] * V[14] -
         s32 * q[12] * V[15];
    ty = -q[2] * V[1] + q[1] * V[2] - s3 * q[5] * V[4] + (s3 * q[4] - q[7]) * V[5] - q[8] * V[6] + q[5] * V[7] +
         q[6] * V[8] - s6 * q[10] * V[9] + (-s52 * q[12] + s6 * q[9]) * V[10] - s52 * q[13] * V[11] +
         (s52 * q[10] - s32 * q[14]) * V[12] + (s52 * q[11] - s32 * q[15]) * V[13] + s32 * q[12] * V[14] +
         s32 * q[13] * V[15];
    tz = -q[3] * V[2] + q[2] * V[3] - q[6] * V[5] + q[5] * V[6] - R(2) * q[8] * V[7] + R(2) * q[7] * V[8] -
         q[11] * V[10] + q[10] * V[11] - R(2) * q[13] * V[12] + R(2) * q[12] * V[13] - R(3) * q[15] * V[14] +
         R(3) * q[14] * V[15];
}

// The charge-charge channel of a pair (scale 1): energy q_a q_b erfc(alpha r)/r
// and the force on b along r.  For two bare charges (no higher moment, no
// polarisability) it is the whole of the quasi-internal-frame interaction
// below.
// kPot: also adds the potentials (energy per unit charge) at a and b to *pa, *pb.
template <class R, bool kPot = false>
MPID_HD void charge_pair(V3<R> dr, R r2, R qa, R qb, const PairParams<R>& P, V3<R>& f, R& energy, R* pa = nullptr,
                         R* pb = nullptr) {
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    const R ar = P.alpha * r;
    const R e = exp_r(-ar * ar);
    const R ec = erfc_r(ar, e);
    const R c = R(kElectric) * qa * qb * rinv;
    if constexpr (kPot) {
        const R g = R(kElectric) * rinv * ec;
        *pa += g * qb;
        *pb += g * qa;
    }
    energy = c * ec;
    f = dr * (c * (ec + ar * R(2.0 / kSqrtPi) * e) * rinv * rinv);
}

template <class R> struct ElecOut {
    V3<R> f;    // force on j (force on i is -f)
    V3<R> ti, tj;
    R energy;
};

// QI frame (formQIRotationMatrix, MPIDReferenceForce.cpp:650-681): z along r,
// x = the lab x axis made orthogonal to z (the lab y axis when r lies on x).
// Written as (s, -zx zy / s, -zx zz / s), s = sqrt(zy^2 + zz^2), which is the
// same vector without the cancellation of 1 - zx^2.  The choice of x matters:
// the reference's induced-induced z torque is not invariant under rotations
// about r (see PairParams::fix_torque).
template <class R> MPID_HD void qi_frame(V3<R> dr, R rinv, V3<R>& ex, V3<R>& ey, V3<R>& ez) {
    ez = dr * rinv;
    const R syz2 = ez.y * ez.y + ez.z * ez.z;
    if (dr.y != R(0) || dr.z != R(0)) {
        const R is = rsqrt_r(syz2);
        ex = mk(syz2 * is, -ez.x * ez.y * is, -ez.x * ez.z * is);
    } else {
        ex = mk(R(0), R(1), R(0));
    }
    ey = cross(ez, ex);
}

// kPot: also adds the pair's potential vectors dE/dm to pot_a and pot_b (20
// stored lab components each, the induced dipoles held fixed); without
// with_cc they leave out the charge-charge channel, as the energy does.
template <class R, bool kPot = false>
MPID_HD ElecOut<R> elec_pair(V3<R> dr, R r2, const R* ma, const R* mb, V3<R> mua, V3<R> mub,
                                                R2<R> tda, R2<R> tdb, bool anisoA, bool anisoB, R ms, R ps,
                                                const PairParams<R>& P, bool with_cc = true, R* pot_a = nullptr,
                                                R* pot_b = nullptr) {
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    V3<R> ex, ey, ez;
    qi_frame(dr, rinv, ex, ey, ez);
    R qa[16], qb[16];
    qi_moments(ma, ex, ey, ez, qa);
    qi_moments(mb, ex, ey, ez, qb);
    R ua[3] = {R(0.5) * dot(ez, mua), R(0.5) * dot(ex, mua), R(0.5) * dot(ey, mua)};
    R ub[3] = {R(0.5) * dot(ez, mub), R(0.5) * dot(ex, mub), R(0.5) * dot(ey, mub)};

    // Ewald and Thole radial factors.  b'_k = 1 + bVec[k] of the reference
    // (computed from erfc, so ms + bVec[k] = (ms - 1) + b'_k loses nothing).
    R ri[9];
    ri[1] = R(kElectric) * rinv;
#pragma unroll
    for (int k = 2; k < 9; k++) ri[k] = ri[k - 1] * rinv;
    R ar[10];
    ar[1] = P.alpha * r;
#pragma unroll
    for (int k = 2; k < 10; k++) ar[k] = ar[k - 1] * ar[1];
    const R eX = exp_r(-ar[2]);
    const R X = R(2.0 / kSqrtPi) * eX;
    const R b2 = erfc_r(ar[1], eX) + ar[1] * X;
    const R b3 = b2 + R(2.0 / 3.0) * ar[3] * X;
    const R b4 = b3 + R(4.0 / 15.0) * ar[5] * X;
    const R b5 = b4 + R(8.0 / 105.0) * ar[7] * X;
    const R m1 = ms - R(1);
    const R mB2 = m1 + b2, mB3 = m1 + b3, mB4 = m1 + b4, mB5 = m1 + b5;
    const R dmp = tda.y * tdb.y;
    const R a = (ps == R(0)) ? tda.x + tdb.x : P.default_thole;
    const R u = fabs(dmp) > R(1e-5) ? div_r(r, dmp) : R(1e10);
    // exp(-au) vanishes for au >= 50; the polynomial argument is clamped there so
    // that its powers stay finite in FP32 (u = 1e10 when a damping factor is 0)
    const R au = fmin(a * u, R(50));
    const R e = au < R(50) ? exp_r(-au) : R(0);
    const R au2 = au * au, au3 = au2 * au, au4 = au3 * au, au5 = au4 * au, au6 = au5 * au;
    // th - 1 for each Thole factor (MPIDReferenceForce.cpp:4509-4523)
    const R base = R(1) + au + R(0.5) * au2;
    const R t_c = -e * base;
    const R t_d0 = -e * (base + au3 * R(0.25));
    const R t_d1 = t_c;
    const R t_q0 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 18.0));
    const R t_q1 = -e * (base + au3 * R(1.0 / 6.0));
    const R t_o0 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 24.0) + au5 * R(1.0 / 120.0));
    const R t_o1 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 30.0));
    const R dt_c = -e * (base + au3 * R(0.25));
    const R dt_d0 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 12.0));
    const R dt_d1 = -e * (base + au3 * R(1.0 / 6.0));
    const R dt_q0 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 24.0) + au5 * R(1.0 / 72.0));
    const R dt_q1 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 24.0));
    const R dt_o0 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 24.0) + au5 * R(1.0 / 120.0) + au6 * R(1.0 / 600.0));
    const R dt_o1 = -e * (base + au3 * R(1.0 / 6.0) + au4 * R(1.0 / 25.0) + au5 * R(1.0 / 150.0));
    // scale*th + bVec[k] = (scale - 1) + scale*(th - 1) + b'_k
    const R p1 = ps - R(1);
    auto sc = [&](R th1) { return p1 + ps * th1; };
    const R X3 = ar[3] * X, X5 = ar[5] * X, X7 = ar[7] * X, X9 = ar[9] * X;

    R Va[16], Vb[16], VaR[16], VbR[16], Vad[3], Vbd[3];
    R eP, dP, eU, dU;
    // C-C (m=0); without with_cc the caller evaluates this channel itself
    // (charge_pair, which it equals for scale 1)
    if (with_cc) {
        eP = ri[1] * (mB2 - ar[1] * X);
        dP = R(-0.5) * mB2 * ri[2];
        Va[0] = eP * qb[0]; Vb[0] = eP * qa[0]; VaR[0] = dP * qb[0]; VbR[0] = dP * qa[0];
    } else {
        Va[0] = Vb[0] = VaR[0] = VbR[0] = R(0);
    }
    // C-D, C-Uind (m=0)
    eP = ri[2] * mB2;
    eU = R(2) * ri[2] * (sc(t_c) + b2);
    dP = -ri[3] * (mB2 + X3);
    dU = R(-4) * ri[3] * (sc(dt_c) + b2 + X3);
    Va[0] += -(eP * qb[1] + eU * ub[0]); Vb[1] = -(eP * qa[0]);
    VaR[0] += -(dP * qb[1] + dU * ub[0]); VbR[1] = -(dP * qa[0]);
    Vbd[0] = -(eU * qa[0]);
    Va[1] = eP * qb[0]; Vb[0] += eP * qa[1] + eU * ua[0];
    VaR[1] = dP * qb[0]; VbR[0] += dP * qa[1] + dU * ua[0];
    Vad[0] = eU * qb[0];
    // D-D, D-Uind (m=0)
    eP = R(-2.0 / 3.0) * ri[3] * (R(3) * mB3 + X3);
    eU = R(-4.0 / 3.0) * ri[3] * (R(3) * (sc(t_d0) + b3) + X3);
    dP = ri[4] * (R(3) * mB3 + R(2) * X5);
    dU = R(2) * ri[4] * (R(6) * (sc(dt_d0) + b3) + R(4) * X5);
    Va[1] += eP * qb[1] + eU * ub[0]; Vb[1] += eP * qa[1] + eU * ua[0];
    VaR[1] += dP * qb[1] + dU * ub[0]; VbR[1] += dP * qa[1] + dU * ua[0];
    Vad[0] += eU * qb[1]; Vbd[0] += eU * qa[1];
    // D-D, D-Uind (m=1)
    eP = ri[3] * (mB3 - R(2.0 / 3.0) * X3);
    eU = R(2) * ri[3] * (sc(t_d1) + b3 - R(2.0 / 3.0) * X3);
    dP = R(-1.5) * ri[4] * mB3;
    dU = R(-6) * ri[4] * (sc(dt_d1) + b3);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        Va[2 + c] = eP * qb[2 + c] + eU * ub[1 + c]; Vb[2 + c] = eP * qa[2 + c] + eU * ua[1 + c];
        VaR[2 + c] = dP * qb[2 + c] + dU * ub[1 + c]; VbR[2 + c] = dP * qa[2 + c] + dU * ua[1 + c];
        Vad[1 + c] = eU * qb[2 + c]; Vbd[1 + c] = eU * qa[2 + c];
    }
    // C-Q, Q-C (m=0)
    eP = mB3 * ri[3];
    dP = R(-1.0 / 3.0) * ri[4] * (R(4.5) * mB3 + R(2) * X5);
    Va[0] += eP * qb[4]; Vb[4] = eP * qa[0]; VaR[0] += dP * qb[4]; VbR[4] = dP * qa[0];
    Va[4] = eP * qb[0]; Vb[0] += eP * qa[4]; VaR[4] = dP * qb[0]; VbR[0] += dP * qa[4];
    // D-Q, Uind-Q (m=0)
    eP = ri[4] * (R(3) * mB3 + R(4.0 / 3.0) * X5);
    eU = R(2) * ri[4] * (R(3) * (sc(t_q0) + b3) + R(4.0 / 3.0) * X5);
    dP = R(-4.0 / 3.0) * ri[5] * (R(4.5) * mB3 + (R(1) + ar[2]) * X5);
    dU = R(-8.0 / 3.0) * ri[5] * (R(9) * (sc(dt_q0) + b3) + R(2) * (R(1) + ar[2]) * X5);
    Va[1] += eP * qb[4]; Vb[4] += eP * qa[1] + eU * ua[0];
    VaR[1] += dP * qb[4]; VbR[4] += dP * qa[1] + dU * ua[0];
    Vad[0] += eU * qb[4];
    Va[4] += -(eP * qb[1] + eU * ub[0]); Vb[1] += -(eP * qa[4]);
    VaR[4] += -(dP * qb[1] + dU * ub[0]); VbR[1] += -(dP * qa[4]);
    Vbd[0] += -(eU * qa[4]);
    // D-Q, Uind-Q (m=1)
    eP = R(-1.7320508075688772) * ri[4] * mB3;
    eU = R(-2 * 1.7320508075688772) * ri[4] * (sc(t_q1) + b3);
    dP = R(4.0 / 1.7320508075688772) * ri[5] * (R(1.5) * mB3 + R(0.5) * X5);
    dU = R(8.0 / 1.7320508075688772) * ri[5] * (R(3) * (sc(dt_q1) + b3) + X5);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        Va[2 + c] += eP * qb[5 + c]; Vb[5 + c] = eP * qa[2 + c] + eU * ua[1 + c];
        VaR[2 + c] += dP * qb[5 + c]; VbR[5 + c] = dP * qa[2 + c] + dU * ua[1 + c];
        Vad[1 + c] += eU * qb[5 + c];
        Va[5 + c] = -(eP * qb[2 + c] + eU * ub[1 + c]); Vb[2 + c] += -(eP * qa[5 + c]);
        VaR[5 + c] = -(dP * qb[2 + c] + dU * ub[1 + c]); VbR[2 + c] += -(dP * qa[5 + c]);
        Vbd[1 + c] += -(eU * qa[5 + c]);
    }
    // Q-Q (m=0, 1, 2)
    eP = ri[5] * (R(6) * mB4 + R(4.0 / 45.0) * (R(-3) + R(10) * ar[2]) * X5);
    dP = R(-1.0 / 9.0) * ri[6] * (R(135) * mB4 + R(4) * (R(1) + R(2) * ar[2]) * X7);
    Va[4] += eP * qb[4]; Vb[4] += eP * qa[4]; VaR[4] += dP * qb[4]; VbR[4] += dP * qa[4];
    eP = R(-4.0 / 15.0) * ri[5] * (R(15) * mB4 + X5);
    dP = ri[6] * (R(10) * mB4 + R(4.0 / 3.0) * X7);
#pragma unroll
    for (int c = 5; c < 7; c++) {
        Va[c] += eP * qb[c]; Vb[c] += eP * qa[c]; VaR[c] += dP * qb[c]; VbR[c] += dP * qa[c];
    }
    eP = ri[5] * (mB4 - R(4.0 / 15.0) * X5);
    dP = R(-2.5) * mB4 * ri[6];
#pragma unroll
    for (int c = 7; c < 9; c++) {
        Va[c] = eP * qb[c]; Vb[c] = eP * qa[c]; VaR[c] = dP * qb[c]; VbR[c] = dP * qa[c];
    }
    // C-O, O-C (m=0)
    eP = ri[4] * (-mB3 - R(4.0 / 15.0) * X5);
    dP = R(2.0 / 15.0) * ri[5] * (R(15) * mB3 + R(2) * (R(2) * X5 + X7));
    Va[0] += eP * qb[9]; Vb[9] = eP * qa[0]; VaR[0] += dP * qb[9]; VbR[9] = dP * qa[0];
    Va[9] = -eP * qb[0]; Vb[0] -= eP * qa[9]; VaR[9] = -dP * qb[0]; VbR[0] -= dP * qa[9];
    // D-O, Uind-O (m=0)
    eP = R(-4) * ri[5] * (mB4 + R(2.0 / 15.0) * X7);
    eU = R(-8) * ri[5] * (sc(t_o0) + b4 + R(2.0 / 15.0) * X7);
    dP = R(2.0 / 15.0) * ri[6] * (R(75) * mB4 + R(4) * (R(1) + ar[2]) * X7);
    dU = R(8.0 / 15.0) * ri[6] * (R(75) * (sc(dt_o0) + b4) + R(4) * (R(1) + ar[2]) * X7);
    Va[1] += eP * qb[9]; Vb[9] += eP * qa[1] + eU * ua[0];
    VaR[1] += dP * qb[9]; VbR[9] += dP * qa[1] + dU * ua[0];
    Va[9] += eP * qb[1] + eU * ub[0]; Vb[1] += eP * qa[9];
    VaR[9] += dP * qb[1] + dU * ub[0]; VbR[1] += dP * qa[9];
    Vad[0] += eU * qb[9]; Vbd[0] += eU * qa[9];
    // D-O, Uind-O (m=1)
    const R s6 = R(2.449489742783178);
    const R c23 = R(0.2 * 0.816496580927726);
    eP = s6 * mB4 * ri[5];
    eU = R(2) * s6 * (sc(t_o1) + b4) * ri[5];
    dP = R(-0.5) * c23 * ri[6] * (R(75) * mB4 + R(8) * X7);
    dU = R(-2) * c23 * ri[6] * (R(75) * (sc(dt_o1) + b4) + R(8) * X7);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        Va[2 + c] += eP * qb[10 + c]; Vb[10 + c] = eP * qa[2 + c] + eU * ua[1 + c];
        VaR[2 + c] += dP * qb[10 + c]; VbR[10 + c] = dP * qa[2 + c] + dU * ua[1 + c];
        Vad[1 + c] += eU * qb[10 + c];
        Va[10 + c] = eP * qb[2 + c] + eU * ub[1 + c]; Vb[2 + c] += eP * qa[10 + c];
        VaR[10 + c] = dP * qb[2 + c] + dU * ub[1 + c]; VbR[2 + c] += dP * qa[10 + c];
        Vbd[1 + c] += eU * qa[10 + c];
    }
    // Q-O, O-Q (m=0)
    eP = ri[6] * (R(-10) * mB4 - R(8.0 / 45.0) * (R(3) + R(2) * ar[2]) * X7);
    dP = R(2.0 / 45.0) * ri[7] * (R(675) * mB4 + R(2) * (R(27) + R(4) * ar[4]) * X7);
    Va[4] += eP * qb[9]; Vb[9] += eP * qa[4]; VaR[4] += dP * qb[9]; VbR[9] += dP * qa[4];
    Va[9] -= eP * qb[4]; Vb[4] -= eP * qa[9]; VaR[9] -= dP * qb[4]; VbR[4] -= dP * qa[9];
    // Q-O, O-Q (m=1)
    eP = R(5 * 1.4142135623730951) * ri[6] * (mB4 + R(8.0 / 75.0) * X7);
    dP = R(-0.5 * 2.8284271247461903 / 15.0) * ri[7] * (R(225) * mB4 + R(8) * (R(2) + ar[2]) * X7);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        Va[5 + c] += eP * qb[10 + c]; Vb[10 + c] += eP * qa[5 + c];
        VaR[5 + c] += dP * qb[10 + c]; VbR[10 + c] += dP * qa[5 + c];
        Va[10 + c] -= eP * qb[5 + c]; Vb[5 + c] -= eP * qa[10 + c];
        VaR[10 + c] -= dP * qb[5 + c]; VbR[5 + c] -= dP * qa[10 + c];
    }
    // Q-O, O-Q (m=2)
    eP = R(-2.23606797749979) * mB4 * ri[6];
    dP = R(2.23606797749979 / 15.0) * ri[7] * (R(45) * mB4 + R(4) * X7);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        Va[7 + c] += eP * qb[12 + c]; Vb[12 + c] = eP * qa[7 + c];
        VaR[7 + c] += dP * qb[12 + c]; VbR[12 + c] = dP * qa[7 + c];
        Va[12 + c] = -eP * qb[7 + c]; Vb[7 + c] -= eP * qa[12 + c];
        VaR[12 + c] = -dP * qb[7 + c]; VbR[7 + c] -= dP * qa[12 + c];
    }
    // O-O (m=0..3)
    eP = ri[7] * (R(-20) * mB5 - R(8.0 / 1575.0) * (R(15) + R(28) * ar[2] + R(28) * ar[4]) * X7);
    dP = R(2.0 / 225.0) * ri[8] * (R(7875) * mB5 + R(4) * (R(41) - R(4) * ar[2] + R(4) * ar[4]) * X9);
    Va[9] += eP * qb[9]; Vb[9] += eP * qa[9]; VaR[9] += dP * qb[9]; VbR[9] += dP * qa[9];
    eP = ri[7] * (R(15) * mB5 + R(8.0 / 525.0) * (R(-5) + R(28) * ar[2]) * X7);
    dP = R(-1.0 / 150.0) * ri[8] * (R(7875) * mB5 + R(32) * (R(3) + R(2) * ar[2]) * X9);
#pragma unroll
    for (int c = 10; c < 12; c++) {
        Va[c] += eP * qb[c]; Vb[c] += eP * qa[c]; VaR[c] += dP * qb[c]; VbR[c] += dP * qa[c];
    }
    eP = ri[7] * (R(-6) * mB5 - R(8.0 / 105.0) * X7);
    dP = R(0.5) * ri[8] * (R(42) * mB5 + R(16.0 / 15.0) * X9);
#pragma unroll
    for (int c = 12; c < 14; c++) {
        Va[c] += eP * qb[c]; Vb[c] += eP * qa[c]; VaR[c] += dP * qb[c]; VbR[c] += dP * qa[c];
    }
    eP = ri[7] * (mB5 - R(8.0 / 105.0) * X7);
    dP = R(-3.5) * mB5 * ri[8];
#pragma unroll
    for (int c = 14; c < 16; c++) {
        Va[c] = eP * qb[c]; Vb[c] = eP * qa[c]; VaR[c] = dP * qb[c]; VbR[c] = dP * qa[c];
    }

    // Va, Vb are dE/dqa, dE/dqb at fixed induced dipoles (the induced terms
    // enter with ua = mua/2 and twice the energy coefficient)
    if constexpr (kPot) {
        qi_potential_to_lab(Va, ex, ey, ez, pot_a);
        qi_potential_to_lab(Vb, ex, ey, ez, pot_b);
    }
    ElecOut<R> o;
    R energy = R(0), fIZ = R(0), fJZ = R(0);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        energy += qa[k] * Va[k] + qb[k] * Vb[k];
        fIZ += qa[k] * VaR[k];
        fJZ += qb[k] * VbR[k];
    }
    o.energy = R(0.5) * energy;
    R EIX, EIY, EIZ, EJX, EJY, EJZ;
    qi_torque_sums(qa, Va, EIX, EIY, EIZ);
    qi_torque_sums(qb, Vb, EJX, EJY, EJZ);
    // induced torque intermediates
    R iEIX = ua[2] * Vad[0] - ua[0] * Vad[2];
    R iEJX = ub[2] * Vbd[0] - ub[0] * Vbd[2];
    R iEIY = ua[0] * Vad[1] - ua[1] * Vad[0];
    R iEJY = ub[0] * Vbd[1] - ub[1] * Vbd[0];
    R iEIZ = ua[1] * Vad[2] - ua[2] * Vad[1];
    R iEJZ = ub[1] * Vbd[2] - ub[2] * Vbd[1];
    if (P.mutual) {
        // induced-induced terms, never scaled (uScale = 1)
        R eC = R(-8.0 / 3.0) * ri[3] * (R(3) * (t_d0 + b3) + X3);
        R dC = R(2) * ri[4] * (R(6) * (dt_d0 + b3) + R(4) * X5);
        iEIX += eC * ua[2] * ub[0];
        iEJX += eC * ub[2] * ua[0];
        iEIY -= eC * ua[1] * ub[0];
        iEJY -= eC * ub[1] * ua[0];
        fIZ += dC * ua[0] * ub[0];
        fJZ += dC * ub[0] * ua[0];
        eC = R(4) * ri[3] * (t_d1 + b3 - R(2.0 / 3.0) * X3);
        dC = R(-6) * ri[4] * (dt_d1 + b3);
        iEIX -= eC * ua[0] * ub[2];
        iEJX -= eC * ub[0] * ua[2];
        iEIY += eC * ua[0] * ub[1];
        iEJY += eC * ub[0] * ua[1];
        iEIZ += eC * ua[1] * ub[2];
        iEJZ += eC * ub[1] * ua[2];
        if (P.fix_torque) {
            // the second half of the rotational derivative about r, absent at
            // MPIDReferenceForce.cpp:4854-4855; with it the torques on
            // anisotropic sites are exact derivatives of the energy
            iEIZ -= eC * ua[2] * ub[1];
            iEJZ -= eC * ub[2] * ua[1];
        }
        fIZ += dC * (ua[1] * ub[1] + ua[2] * ub[2]);
        fJZ += dC * (ub[1] * ua[1] + ub[2] * ua[2]);
    }
    const R qfx = rinv * (EIY + EJY + iEIY + iEJY);
    const R qfy = -rinv * (EIX + EJX + iEIX + iEJX);
    const R qfz = -(fJZ + fIZ);
    R tix = -EIX, tiy = -EIY, tiz = -EIZ, tjx = -EJX, tjy = -EJY, tjz = -EJZ;
    if (anisoA) { tix -= iEIX; tiy -= iEIY; tiz -= iEIZ; }
    if (anisoB) { tjx -= iEJX; tjy -= iEJY; tjz -= iEJZ; }
    o.f = ex * qfx + ey * qfy + ez * qfz;
    o.ti = ex * tix + ey * tiy + ez * tiz;
    o.tj = ex * tjx + ey * tjy + ez * tjz;
    return o;
}

template <class R> struct ElecBareOut {
    V3<R> f;   // force on b (force on a is -f)
    V3<R> ta;  // torque on a (b, a point charge, takes none)
    R energy;
};

// elec_pair with with_cc = false for a pair whose site b is a bare charge:
// qb[k] = 0 for k >= 1 and no induced dipole (no polarisability).  Every
// term of elec_pair with a factor qb[k >= 1] or ub vanishes, and so do b's
// torque and all induced-induced terms; what remains is b's charge against
// a's dipole, quadrupole, octopole (the m = 0 components along r) and a's
// induced dipole.  Same formulas, same order of operations, for those terms.
// kPot: also adds a's potential vector (without its charge-charge channel)
// to pot_a and b's potential (charge component only) to *pot_b0.
template <class R, bool kPot = false>
MPID_HD ElecBareOut<R> elec_pair_bare(V3<R> dr, R r2, const R* ma, R qb, V3<R> mua, R2<R> tda, R2<R> tdb,
                                       bool anisoA, R ms, R ps, const PairParams<R>& P, R* pot_a = nullptr,
                                       R* pot_b0 = nullptr) {
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    V3<R> ex, ey, ez;
    qi_frame(dr, rinv, ex, ey, ez);
    R qa[16];
    qi_moments(ma, ex, ey, ez, qa);
    const R ua0 = R(0.5) * dot(ez, mua), ua1 = R(0.5) * dot(ex, mua), ua2 = R(0.5) * dot(ey, mua);

    R ri[6];
    ri[1] = R(kElectric) * rinv;
#pragma unroll
    for (int k = 2; k < 6; k++) ri[k] = ri[k - 1] * rinv;
    R ar[8];
    ar[1] = P.alpha * r;
#pragma unroll
    for (int k = 2; k < 8; k++) ar[k] = ar[k - 1] * ar[1];
    const R eX = exp_r(-ar[2]);
    const R X = R(2.0 / kSqrtPi) * eX;
    const R b2 = erfc_r(ar[1], eX) + ar[1] * X;
    const R b3 = b2 + R(2.0 / 3.0) * ar[3] * X;
    const R m1 = ms - R(1);
    const R mB2 = m1 + b2, mB3 = m1 + b3;
    const R dmp = tda.y * tdb.y;
    const R a = (ps == R(0)) ? tda.x + tdb.x : P.default_thole;
    const R u = fabs(dmp) > R(1e-5) ? div_r(r, dmp) : R(1e10);
    const R au = fmin(a * u, R(50));
    const R e = au < R(50) ? exp_r(-au) : R(0);
    const R au2 = au * au, au3 = au2 * au;
    const R base = R(1) + au + R(0.5) * au2;
    const R t_c = -e * base;
    const R dt_c = -e * (base + au3 * R(0.25));
    const R p1 = ps - R(1);
    const R X3 = ar[3] * X, X5 = ar[5] * X, X7 = ar[7] * X;

    // C-D, C-Uind (m=0)
    R eP = ri[2] * mB2;
    const R eU = R(2) * ri[2] * (p1 + ps * t_c + b2);
    R dP = -ri[3] * (mB2 + X3);
    const R dU = R(-4) * ri[3] * (p1 + ps * dt_c + b2 + X3);
    const R Va1 = eP * qb, VaR1 = dP * qb, Vad0 = eU * qb;
    R Vb0 = eP * qa[1] + eU * ua0;
    R VbR0 = dP * qa[1] + dU * ua0;
    // C-Q (m=0)
    eP = mB3 * ri[3];
    dP = R(-1.0 / 3.0) * ri[4] * (R(4.5) * mB3 + R(2) * X5);
    const R Va4 = eP * qb, VaR4 = dP * qb;
    Vb0 += eP * qa[4];
    VbR0 += dP * qa[4];
    // C-O (m=0)
    eP = ri[4] * (-mB3 - R(4.0 / 15.0) * X5);
    dP = R(2.0 / 15.0) * ri[5] * (R(15) * mB3 + R(2) * (R(2) * X5 + X7));
    const R Va9 = -eP * qb, VaR9 = -dP * qb;
    Vb0 -= eP * qa[9];
    VbR0 -= dP * qa[9];

    if constexpr (kPot) {
        qi_potential_to_lab_m0(Va1, Va4, Va9, ez, pot_a);
        *pot_b0 += Vb0;
    }
    ElecBareOut<R> o;
    o.energy = R(0.5) * (qa[1] * Va1 + qa[4] * Va4 + qa[9] * Va9 + qb * Vb0);
    const R fIZ = qa[1] * VaR1 + qa[4] * VaR4 + qa[9] * VaR9, fJZ = qb * VbR0;
    // qi_torque_sums with V nonzero at 1, 4 and 9 only (the z sum vanishes)
    const R EIX = qa[3] * Va1 + R(1.7320508075688772) * qa[6] * Va4 + R(2.449489742783178) * qa[11] * Va9;
    const R EIY = -qa[2] * Va1 - R(1.7320508075688772) * qa[5] * Va4 - R(2.449489742783178) * qa[10] * Va9;
    const R iEIX = ua2 * Vad0, iEIY = -ua1 * Vad0;
    const R qfx = rinv * (EIY + iEIY);
    const R qfy = -rinv * (EIX + iEIX);
    const R qfz = -(fJZ + fIZ);
    R tix = -EIX, tiy = -EIY;
    if (anisoA) { tix -= iEIX; tiy -= iEIY; }
    o.f = ex * qfx + ey * qfy + ez * qfz;
    o.ta = ex * tix + ey * tiy;
    return o;
}

// The pair of elec_pair_bare at scale factors 1 (every pair of the tile
// kernel), in Cartesian form.  b's charge sits in the potential of a's
// moments, phi(dr) = B1 (d.dr) + B2 (dr.Q.dr) + B3 (O:dr dr dr) with
// dr = r_b - r_a and B_k the Ewald-screened radial functions (B_k' =
// -r B_(k+1)); the charge-charge term is left to the caller, as there.  b is
// not polarisable, so no Thole damping enters.  a's induced dipole adds to
// d: half of it in the energy (the reference's polarization energy), all of
// it in the force, and in the torque on a only when a is anisotropic (the
// reference's convention).  Energy q_b phi, force on b -q_b grad phi, torque
// on a -q_b (B1 d x dr + 2 B2 (Q.dr) x dr + 3 B3 (O:dr dr) x dr).  Same
// quantities as elec_pair_bare, in about a third of its arithmetic.
// kPot: adds dE/dm of a's stored components (the charge one excepted) to
// pot_a and b's potential to *pot_b0, as elec_pair_bare<R, true>.
template <class R, bool kPot = false>
MPID_HD ElecBareOut<R> elec_pair_bare_cart(V3<R> dr, R r2, const R* ma, R qb, V3<R> mua, bool anisoA,
                                           const PairParams<R>& P, R* pot_a = nullptr, R* pot_b0 = nullptr) {
    const R rinv = rsqrt_r(r2);
    const R r = r2 * rinv;
    const R rinv2 = rinv * rinv;
    const R ar = P.alpha * r;
    const R ex2 = exp_r(-ar * ar);
    const R B0 = erfc_r(ar, ex2) * rinv;
    const R B1 = (B0 + P.aln[1] * ex2) * rinv2;
    const R B2 = (R(3) * B1 + P.aln[2] * ex2) * rinv2;
    const R B3 = (R(5) * B2 + P.aln[3] * ex2) * rinv2;
    const R B4 = (R(7) * B3 + P.aln[4] * ex2) * rinv2;
    const V3<R> d = mk(ma[1], ma[2], ma[3]);
    const V3<R> qr = quad_dot(ma + 4, dr);  // Q.dr
    R M[6];
    oct_dot(ma + 10, dr, M);
    const V3<R> orr = quad_dot(M, dr);  // O:dr dr
    const R dd = dot(d, dr), du = dot(mua, dr), qrr = dot(qr, dr), orrr = dot(orr, dr);
    const R ce = R(kElectric) * qb;
    ElecBareOut<R> o;
    o.energy = ce * (B1 * (dd + R(0.5) * du) + B2 * qrr + B3 * orrr);
    // grad[B1 d.r] = B1 d - B2 r (d.r); grad[B2 r.Q.r] = 2 B2 Q.r - B3 r (r.Q.r);
    // grad[B3 O:rrr] = 3 B3 O:rr - B4 r (O:rrr)
    const V3<R> dfv = d + mua;
    const R df = dd + du;
    const V3<R> g = dfv * B1 + qr * (R(2) * B2) + orr * (R(3) * B3) - dr * (B2 * df + B3 * qrr + B4 * orrr);
    o.f = g * (-ce);
    const V3<R> td = anisoA ? dfv : d;
    o.ta = (cross(td, dr) * B1 + cross(qr, dr) * (R(2) * B2) + cross(orr, dr) * (R(3) * B3)) * (-ce);
    if constexpr (kPot) {
        const R c1 = ce * B1, c2 = ce * B2, c3 = ce * B3;
        pot_a[1] += c1 * dr.x;
        pot_a[2] += c1 * dr.y;
        pot_a[3] += c1 * dr.z;
        sym2_add(c2, dr, dr, pot_a + 4);
        sym3_add(c3, dr, dr, dr, pot_a + 10);
        *pot_b0 += R(kElectric) * (B1 * df + B2 * qrr + B3 * orrr);
    }
    return o;
}

}  // namespace mpid
