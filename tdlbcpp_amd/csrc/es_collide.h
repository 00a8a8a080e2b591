// Per-cell arithmetic of the Eggels-Somers scheme, host and device.
//
// moments:   tdlbcpp/src/CollisionEgglesSomers.hpp:185-256
// collision: tdlbcpp/src/CollisionEgglesSomers.hpp:26-181 (+ velocity QVec.hpp:125-134)
//
// T is the type of the stored values, P the type the reference's double
// literals promote to.  With P = double (the reference arithmetic) the reference
// is compiled with T = float and double literals, so several sub-expressions
// are evaluated in double and rounded back to T; every expression below spells
// those promotions out as (D) casts and keeps the reference's left-to-right order
// of terms, so that results are identical to the reference.  With P = float the
// same expressions run entirely in FP32.  All code that includes this header
// is compiled with -ffp-contract=off; es_collide_dev, the float-only collision
// of the 16-bit storage formats, alone turns contraction on for its body.
//
// One deliberate rewrite: the reference divides each of the 18 relaxed moments
// by 24.0 in double and rounds to float.  For T = float we multiply by the
// double 1/24 instead; (float)((double)x * (1/24.0)) equals
// (float)((double)x / 24.0) for every non-NaN float x (checked exhaustively over
// all 2^32 bit patterns by tests/test_device_arith.py).
#pragma once

#include <cmath>
#include "lattice.h"

namespace tdlb {

template <typename T>
struct CollideParams {
    T nu;      // kinematic viscosity (FlowParams::nu after calcNuAndRe_m)
    T cs0;     // Smagorinsky constant
    T g3;      // third-order relaxation
    T b0, c0;  // 1/(1+6nu), 1-6nu in T, computed once per collision as the reference does
    int use_les;
};

template <typename T, typename D = double>
TDLB_HD inline CollideParams<T> make_collide_params(T nu, T cs0, T g3, int use_les) {
    CollideParams<T> p;
    p.nu = nu;
    p.cs0 = cs0;
    p.g3 = g3;
    p.use_les = use_les;
    p.b0 = (T)(D(1.0) / (D(1.0) + (D)(T(6) * nu)));
    p.c0 = (T)(D(1.0) - (D)(T(6) * nu));
    return p;
}

// sqrt(fabs(x)) in the precision of x (the reference's T = float build uses the
// float overloads; both are correctly rounded on the host and on gfx950).
TDLB_HD inline float sqrt_abs(float x) { return sqrtf(fabsf(x)); }
TDLB_HD inline double sqrt_abs(double x) { return sqrt(fabs(x)); }

template <typename T, typename D>
struct Div24 {
    TDLB_HD static T apply(T x) { return (T)((D)x / D(24.0)); }
};
template <>
struct Div24<float, double> {
    TDLB_HD static float apply(float x) { return (float)((double)x * (1.0 / 24.0)); }
};
// Float-only arithmetic (the 16-bit storage formats): the product with
// fl(1/24) is corrected by its exact residual x - 24 q, which gives the
// correctly rounded quotient; the bare product is 3e-8 too large on average
// and grew the mass of every cell by that much per step.
template <>
struct Div24<float, float> {
    TDLB_HD static float apply(float x) {
        const float y = 1.0f / 24.0f;
        const float q = x * y;
        return fmaf(fmaf(-q, 24.0f, x), y, q);
    }
};

// Eighteen moments of eighteen populations; same term order as the reference.
template <typename T, typename D = double>
TDLB_HD inline void es_moments(const T* f, T* m) {
    const T two = T(2), three = T(3);
    m[M01] = f[Q01] + f[Q03] + f[Q02] + f[Q04] + f[Q05] + f[Q06] + f[Q07] + f[Q14] + f[Q08] + f[Q13] + f[Q09]
           + f[Q16] + f[Q10] + f[Q15] + f[Q11] + f[Q18] + f[Q12] + f[Q17];
    m[M02] = f[Q01] - f[Q02] + f[Q07] - f[Q14] - f[Q08] + f[Q13] + f[Q09] - f[Q16] - f[Q10] + f[Q15];
    m[M03] = f[Q03] - f[Q04] + f[Q07] + f[Q14] - f[Q08] - f[Q13] + f[Q11] - f[Q18] - f[Q12] + f[Q17];
    m[M04] = f[Q05] - f[Q06] + f[Q09] + f[Q16] - f[Q10] - f[Q15] + f[Q11] + f[Q18] - f[Q12] - f[Q17];
    m[M05] = -f[Q03] - f[Q04] - f[Q05] - f[Q06] + f[Q07] + f[Q14] + f[Q08] + f[Q13] + f[Q09] + f[Q16] + f[Q10]
           + f[Q15];
    m[M06] = f[Q07] - f[Q14] + f[Q08] - f[Q13];
    m[M07] = -f[Q01] - f[Q02] - f[Q05] - f[Q06] + f[Q07] + f[Q14] + f[Q08] + f[Q13] + f[Q11] + f[Q18] + f[Q12]
           + f[Q17];
    m[M08] = f[Q09] - f[Q16] + f[Q10] - f[Q15];
    m[M09] = f[Q11] - f[Q18] + f[Q12] - f[Q17];
    m[M10] = -f[Q01] - f[Q03] - f[Q02] - f[Q04] + f[Q09] + f[Q16] + f[Q10] + f[Q15] + f[Q11] + f[Q18] + f[Q12]
           + f[Q17];
    m[M11] = -f[Q01] + f[Q02] + two * f[Q07] - two * f[Q14] - two * f[Q08] + two * f[Q13] - f[Q09] + f[Q16]
           + f[Q10] - f[Q15];
    m[M12] = -f[Q03] + f[Q04] + two * f[Q07] + two * f[Q14] - two * f[Q08] - two * f[Q13] - f[Q11] + f[Q18]
           + f[Q12] - f[Q17];
    m[M13] = -three * f[Q01] + three * f[Q02] + three * f[Q09] - three * f[Q16] - three * f[Q10]
           + three * f[Q15];
    m[M14] = -three * f[Q03] + three * f[Q04] + three * f[Q11] - three * f[Q18] - three * f[Q12]
           + three * f[Q17];
    m[M15] = -two * f[Q05] + two * f[Q06] + f[Q09] + f[Q16] - f[Q10] - f[Q15] + f[Q11] + f[Q18] - f[Q12]
           - f[Q17];
    m[M16] = -three * f[Q09] - three * f[Q16] + three * f[Q10] + three * f[Q15] + three * f[Q11]
           + three * f[Q18] - three * f[Q12] - three * f[Q17];
    // fourth order: double literals 0.5 / 1.5 make these sums double
    const D h = D(0.5), o = D(1.5);
    m[M17] = (T)(h * (D)f[Q01] + h * (D)f[Q03] + h * (D)f[Q02] + h * (D)f[Q04] - (D)f[Q05] - (D)f[Q06]
                 - (D)f[Q07] - (D)f[Q14] - (D)f[Q08] - (D)f[Q13] + h * (D)f[Q09] + h * (D)f[Q16]
                 + h * (D)f[Q10] + h * (D)f[Q15] + h * (D)f[Q11] + h * (D)f[Q18] + h * (D)f[Q12]
                 + h * (D)f[Q17]);
    m[M18] = (T)(o * (D)f[Q01] - o * (D)f[Q03] + o * (D)f[Q02] - o * (D)f[Q04] - o * (D)f[Q09]
                 - o * (D)f[Q16] - o * (D)f[Q10] - o * (D)f[Q15] + o * (D)f[Q11] + o * (D)f[Q18]
                 + o * (D)f[Q12] + o * (D)f[Q17]);
}

// Only density and momentum (what IB forcing reads), same arithmetic as es_moments.
template <typename T>
TDLB_HD inline void es_rho_j(const T* f, T& rho, T& jx, T& jy, T& jz) {
    rho = f[Q01] + f[Q03] + f[Q02] + f[Q04] + f[Q05] + f[Q06] + f[Q07] + f[Q14] + f[Q08] + f[Q13] + f[Q09]
        + f[Q16] + f[Q10] + f[Q15] + f[Q11] + f[Q18] + f[Q12] + f[Q17];
    jx = f[Q01] - f[Q02] + f[Q07] - f[Q14] - f[Q08] + f[Q13] + f[Q09] - f[Q16] - f[Q10] + f[Q15];
    jy = f[Q03] - f[Q04] + f[Q07] + f[Q14] - f[Q08] - f[Q13] + f[Q11] - f[Q18] - f[Q12] + f[Q17];
    jz = f[Q05] - f[Q06] + f[Q09] + f[Q16] - f[Q10] - f[Q15] + f[Q11] + f[Q18] - f[Q12] - f[Q17];
}

// Collision of one cell: m (moments) in, p (post-collision populations) out.
// nuc is the cell's eddy viscosity: read (previous step) and written with LES.
template <typename T, typename D = double>
TDLB_HD inline void es_collide(const T* m, T fx, T fy, T fz, T& nuc, const CollideParams<T>& cp, T* p) {
    const D one = D(1.0), half = D(0.5), two = D(2.0);
    const D inv_rho = one / (D)m[M01];
    const T ux = (T)(inv_rho * ((D)m[M02] + half * (D)fx));
    const T uy = (T)(inv_rho * ((D)m[M03] + half * (D)fy));
    const T uz = (T)(inv_rho * ((D)m[M04] + half * (D)fz));

    T b = cp.b0, c = cp.c0;
    if (cp.use_les) {
        const T fct = (T)(D(3.0) / ((D)m[M01] * (one + D(6.0) * (D)(nuc + cp.nu))));
        const D hfx = half * (D)fx, hfy = half * (D)fy, hfz = half * (D)fz;
        const T dudx = (T)((D)fct * (((D)m[M02] + hfx * (D)ux) - (D)m[M05]));
        const T dvdy = (T)((D)fct * (((D)m[M03] + hfy * (D)uy) - (D)m[M07]));
        const T dwdz = (T)((D)fct * (((D)m[M04] + hfz * (D)uz) - (D)m[M10]));
        const T divv = dudx + dvdy + dwdz;
        const T fct2 = T(2) * fct;
        const T sxy = (T)((D)fct2 * (((D)m[M03] + hfy * (D)ux) - (D)m[M06]));
        const T sxz = (T)((D)fct2 * (((D)m[M04] + hfz * (D)ux) - (D)m[M08]));
        const T syz = (T)((D)fct2 * (((D)m[M04] + hfz * (D)uy) - (D)m[M09]));
        const T sh_t = T(2) * dudx * dudx + T(2) * dvdy * dvdy + T(2) * dwdz * dwdz + sxy * sxy + sxz * sxz
                     + syz * syz;
        const T sh = (T)((D)sh_t - (D(2.0) / D(3.0)) * (D)divv * (D)divv);
        nuc = cp.cs0 * cp.cs0 * sqrt_abs(sh);
        const T nut = nuc + cp.nu;
        b = (T)(one / (one + (D)(T(6) * nut)));
        c = (T)(one - (D)(T(6) * nut));
    }

    const D jx2 = two * ((D)m[M02] + half * (D)fx);
    const D jy2 = two * ((D)m[M03] + half * (D)fy);
    const D jz2 = two * ((D)m[M04] + half * (D)fz);
    const D bd = (D)b;

    T a[kQ];
    a[M01] = m[M01];
    a[M02] = m[M02] + fx;
    a[M03] = m[M03] + fy;
    a[M04] = m[M04] + fz;
    a[M05] = (T)((jx2 * (D)ux - (D)(m[M05] * c)) * bd);
    a[M06] = (T)((jx2 * (D)uy - (D)(m[M06] * c)) * bd);
    a[M07] = (T)((jy2 * (D)uy - (D)(m[M07] * c)) * bd);
    a[M08] = (T)((jx2 * (D)uz - (D)(m[M08] * c)) * bd);
    a[M09] = (T)((jy2 * (D)uz - (D)(m[M09] * c)) * bd);
    a[M10] = (T)((jz2 * (D)uz - (D)(m[M10] * c)) * bd);
    const T ng3 = -cp.g3;
#pragma unroll
    for (int l = M11; l <= M16; l++) a[l] = ng3 * m[l];
    a[M17] = T(0.0);
    a[M18] = T(0.0);
#pragma unroll
    for (int l = 0; l < kQ; l++) a[l] = Div24<T, D>::apply(a[l]);

    const T t2 = T(2), t3 = T(3), t4 = T(4), t6 = T(6);
    const D o = D(1.5);
    // face populations: integer coefficients, all in T
    p[Q01] = t2 * a[M01] + t4 * a[M02] + t3 * a[M05] - t3 * a[M07] - t3 * a[M10] - t2 * a[M11] - t2 * a[M13]
           + t2 * a[M17] + t2 * a[M18];
    p[Q02] = t2 * a[M01] - t4 * a[M02] + t3 * a[M05] - t3 * a[M07] - t3 * a[M10] + t2 * a[M11] + t2 * a[M13]
           + t2 * a[M17] + t2 * a[M18];
    p[Q03] = t2 * a[M01] + t4 * a[M03] - t3 * a[M05] + t3 * a[M07] - t3 * a[M10] - t2 * a[M12] - t2 * a[M14]
           + t2 * a[M17] - t2 * a[M18];
    p[Q04] = t2 * a[M01] - t4 * a[M03] - t3 * a[M05] + t3 * a[M07] - t3 * a[M10] + t2 * a[M12] + t2 * a[M14]
           + t2 * a[M17] - t2 * a[M18];
    p[Q05] = t2 * a[M01] + t4 * a[M04] - t3 * a[M05] - t3 * a[M07] + t3 * a[M10] - t4 * a[M15] - t4 * a[M17];
    p[Q06] = t2 * a[M01] - t4 * a[M04] - t3 * a[M05] - t3 * a[M07] + t3 * a[M10] + t4 * a[M15] - t4 * a[M17];
    // edge populations: the first 1.5 coefficient turns the rest of the sum into double
    p[Q07] = (T)((D)(a[M01] + t2 * a[M02] + t2 * a[M03]) + o * (D)a[M05] + (D)(t6 * a[M06]) + o * (D)a[M07]
                 - o * (D)a[M10] + (D)(t2 * a[M11]) + (D)(t2 * a[M12]) - (D)(t2 * a[M17]));
    p[Q14] = (T)((D)(a[M01] - t2 * a[M02] + t2 * a[M03]) + o * (D)a[M05] - (D)(t6 * a[M06]) + o * (D)a[M07]
                 - o * (D)a[M10] - (D)(t2 * a[M11]) + (D)(t2 * a[M12]) - (D)(t2 * a[M17]));
    p[Q08] = (T)((D)(a[M01] - t2 * a[M02] - t2 * a[M03]) + o * (D)a[M05] + (D)(t6 * a[M06]) + o * (D)a[M07]
                 - o * (D)a[M10] - (D)(t2 * a[M11]) - (D)(t2 * a[M12]) - (D)(t2 * a[M17]));
    p[Q13] = (T)((D)(a[M01] + t2 * a[M02] - t2 * a[M03]) + o * (D)a[M05] - (D)(t6 * a[M06]) + o * (D)a[M07]
                 - o * (D)a[M10] + (D)(t2 * a[M11]) - (D)(t2 * a[M12]) - (D)(t2 * a[M17]));
    p[Q09] = (T)((D)(a[M01] + t2 * a[M02] + t2 * a[M04]) + o * (D)a[M05] - o * (D)a[M07] + (D)(t6 * a[M08])
                 + o * (D)a[M10] - (D)a[M11] + (D)a[M13] + (D)a[M15] - (D)a[M16] + (D)a[M17] - (D)a[M18]);
    p[Q16] = (T)((D)(a[M01] - t2 * a[M02] + t2 * a[M04]) + o * (D)a[M05] - o * (D)a[M07] - (D)(t6 * a[M08])
                 + o * (D)a[M10] + (D)a[M11] - (D)a[M13] + (D)a[M15] - (D)a[M16] + (D)a[M17] - (D)a[M18]);
    p[Q10] = (T)((D)(a[M01] - t2 * a[M02] - t2 * a[M04]) + o * (D)a[M05] - o * (D)a[M07] + (D)(t6 * a[M08])
                 + o * (D)a[M10] + (D)a[M11] - (D)a[M13] - (D)a[M15] + (D)a[M16] + (D)a[M17] - (D)a[M18]);
    p[Q15] = (T)((D)(a[M01] + t2 * a[M02] - t2 * a[M04]) + o * (D)a[M05] - o * (D)a[M07] - (D)(t6 * a[M08])
                 + o * (D)a[M10] - (D)a[M11] + (D)a[M13] - (D)a[M15] + (D)a[M16] + (D)a[M17] - (D)a[M18]);
    p[Q11] = (T)((D)(a[M01] + t2 * a[M03] + t2 * a[M04]) - o * (D)a[M05] + o * (D)a[M07] + (D)(t6 * a[M09])
                 + o * (D)a[M10] - (D)a[M12] + (D)a[M14] + (D)a[M15] + (D)a[M16] + (D)a[M17] + (D)a[M18]);
    p[Q18] = (T)((D)(a[M01] - t2 * a[M03] + t2 * a[M04]) - o * (D)a[M05] + o * (D)a[M07] - (D)(t6 * a[M09])
                 + o * (D)a[M10] + (D)a[M12] - (D)a[M14] + (D)a[M15] + (D)a[M16] + (D)a[M17] + (D)a[M18]);
    p[Q12] = (T)((D)(a[M01] - t2 * a[M03] - t2 * a[M04]) - o * (D)a[M05] + o * (D)a[M07] + (D)(t6 * a[M09])
                 + o * (D)a[M10] + (D)a[M12] - (D)a[M14] - (D)a[M15] - (D)a[M16] + (D)a[M17] + (D)a[M18]);
    p[Q17] = (T)((D)(a[M01] + t2 * a[M03] - t2 * a[M04]) - o * (D)a[M05] + o * (D)a[M07] - (D)(t6 * a[M09])
                 + o * (D)a[M10] - (D)a[M12] + (D)a[M14] - (D)a[M15] - (D)a[M16] + (D)a[M17] + (D)a[M18]);
}

// Moments and collision of one cell in float only, for the 16-bit storage
// formats, which carry no bit-for-bit obligation.  It works on deviations from a
// rest state B (B_d = rho_b w_d): f[d] = f_d - B_d in, p[d] = p_d - B_d out (p
// may alias f).  B's moments are (rho_b, 0, ..., 0) and B relaxes to itself,
// so the deviations obey the same linear maps as es_moments / es_collide with
// the density moment replaced by rho - rho_b; the density rho itself enters
// only the velocity and the LES factor.  Every rounding is then relative to
// the deviations (|f - B| ~ 1e-3 .. 5e-2) rather than to the populations
// (~0.04 .. 0.1), which keeps the mass of a cell to ~1e-10 per step where the
// same arithmetic on whole populations drifted by 1e-8.  The maps are factored
// through the sums and differences of the nine opposite pairs, M17/M18 are not
// formed (their relaxed values are zero), multiply-adds are fused.  nuc as in
// es_collide.  Ops supplies the three reciprocals and the square root of the
// LES: exactly rounded by default (ExactOps); a device caller may pass the
// hardware's 1-ulp instructions, whose error is far below the 16-bit rounding.
struct ExactOps {
    TDLB_HD static float rcp(float x) { return 1.0f / x; }
    TDLB_HD static float sqrt(float x) { return sqrtf(x); }
};
template <typename Ops = ExactOps>
TDLB_HD inline void es_collide_dev(const float* f, float rho_b, float fx, float fy, float fz, float& nuc,
                                   const CollideParams<float>& cp, float* p) {
#pragma clang fp contract(fast)
    // sums (s) and differences (a) of the nine opposite pairs
    const float sx = f[Q01] + f[Q02], ax = f[Q01] - f[Q02];
    const float sy = f[Q03] + f[Q04], ay = f[Q03] - f[Q04];
    const float sz = f[Q05] + f[Q06], az = f[Q05] - f[Q06];
    const float sxy = f[Q07] + f[Q08], axy = f[Q07] - f[Q08];    // +x+y / -x-y
    const float sxz = f[Q09] + f[Q10], axz = f[Q09] - f[Q10];    // +x+z / -x-z
    const float syz = f[Q11] + f[Q12], ayz = f[Q11] - f[Q12];    // +y+z / -y-z
    const float sxmy = f[Q13] + f[Q14], axmy = f[Q13] - f[Q14];  // +x-y / -x+y
    const float sxmz = f[Q15] + f[Q16], axmz = f[Q15] - f[Q16];  // +x-z / -x+z
    const float symz = f[Q17] + f[Q18], aymz = f[Q17] - f[Q18];  // +y-z / -y+z

    const float m01 = ((sx + sy) + (sz + sxy)) + ((sxz + syz) + (sxmy + sxmz)) + symz;  // rho - rho_b
    const float m02 = ax + (axy + axmy) + (axz + axmz);
    const float m03 = ay + (axy - axmy) + (ayz + aymz);
    const float m04 = az + (axz - axmz) + (ayz - aymz);
    const float exy = sxy + sxmy, exz = sxz + sxmz, eyz = syz + symz;
    const float m05 = (exy + exz) - (sy + sz);
    const float m06 = sxy - sxmy;
    const float m07 = (exy + eyz) - (sx + sz);
    const float m08 = sxz - sxmz;
    const float m09 = syz - symz;
    const float m10 = (exz + eyz) - (sx + sy);
    const float m11 = 2.0f * (axy + axmy) - (ax + axz + axmz);
    const float m12 = 2.0f * (axy - axmy) - (ay + ayz + aymz);
    const float m13 = 3.0f * ((axz + axmz) - ax);
    const float m14 = 3.0f * ((ayz + aymz) - ay);
    const float m15 = (axz - axmz) + (ayz - aymz) - 2.0f * az;
    const float m16 = 3.0f * ((axmz - axz) + (ayz - aymz));

    const float rho = rho_b + m01;
    const float inv_rho = Ops::rcp(rho);
    const float ux = inv_rho * (m02 + 0.5f * fx);
    const float uy = inv_rho * (m03 + 0.5f * fy);
    const float uz = inv_rho * (m04 + 0.5f * fz);
    float b = cp.b0, c = cp.c0;
    if (cp.use_les) {
        const float fct = 3.0f * Ops::rcp(rho * (1.0f + 6.0f * (nuc + cp.nu)));
        const float hfx = 0.5f * fx, hfy = 0.5f * fy, hfz = 0.5f * fz;
        const float dudx = fct * ((m02 + hfx * ux) - m05);
        const float dvdy = fct * ((m03 + hfy * uy) - m07);
        const float dwdz = fct * ((m04 + hfz * uz) - m10);
        const float divv = dudx + dvdy + dwdz;
        const float fct2 = 2.0f * fct;
        const float sh_xy = fct2 * ((m03 + hfy * ux) - m06);
        const float sh_xz = fct2 * ((m04 + hfz * ux) - m08);
        const float sh_yz = fct2 * ((m04 + hfz * uy) - m09);
        const float sh = 2.0f * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + sh_xy * sh_xy + sh_xz * sh_xz +
                         sh_yz * sh_yz - (2.0f / 3.0f) * divv * divv;
        nuc = cp.cs0 * cp.cs0 * Ops::sqrt(fabsf(sh));
        const float nut6 = 6.0f * (nuc + cp.nu);
        b = Ops::rcp(1.0f + nut6);
        c = 1.0f - nut6;
    }
    const float jx2 = 2.0f * m02 + fx, jy2 = 2.0f * m03 + fy, jz2 = 2.0f * m04 + fz;
    // relaxed moments / 24; b / 24 and -g3 / 24 are folded into one factor each
    const float b24 = b * (1.0f / 24.0f), g24 = cp.g3 * (-1.0f / 24.0f), r24 = 1.0f / 24.0f;
    const float a01 = m01 * r24;
    const float a02 = (m02 + fx) * r24, a03 = (m03 + fy) * r24, a04 = (m04 + fz) * r24;
    const float a05 = (jx2 * ux - m05 * c) * b24, a06 = (jx2 * uy - m06 * c) * b24;
    const float a07 = (jy2 * uy - m07 * c) * b24, a08 = (jx2 * uz - m08 * c) * b24;
    const float a09 = (jy2 * uz - m09 * c) * b24, a10 = (jz2 * uz - m10 * c) * b24;
    const float a11 = g24 * m11, a12 = g24 * m12, a13 = g24 * m13;
    const float a14 = g24 * m14, a15 = g24 * m15, a16 = g24 * m16;

    // faces: p = C +- D
    const float two01 = 2.0f * a01;
    const float cx = two01 + 3.0f * ((a05 - a07) - a10), dx = 4.0f * a02 - 2.0f * (a11 + a13);
    const float cy = two01 + 3.0f * ((a07 - a05) - a10), dy = 4.0f * a03 - 2.0f * (a12 + a14);
    const float cz = two01 + 3.0f * (a10 - (a05 + a07)), dz = 4.0f * (a04 - a15);
    p[Q01] = cx + dx;
    p[Q02] = cx - dx;
    p[Q03] = cy + dy;
    p[Q04] = cy - dy;
    p[Q05] = cz + dz;
    p[Q06] = cz - dz;
    // edges: p = E +- 6 a_off +- G / H
    const float exy2 = a01 + 1.5f * ((a05 + a07) - a10);
    const float exz2 = a01 + 1.5f * ((a05 + a10) - a07);
    const float eyz2 = a01 + 1.5f * ((a07 + a10) - a05);
    const float gxy = 2.0f * ((a02 + a03) + (a11 + a12)), hxy = 2.0f * ((a02 - a03) + (a11 - a12));
    const float gxz = 2.0f * (a02 + a04) + ((a13 - a11) + (a15 - a16));
    const float hxz = 2.0f * (a02 - a04) + ((a13 - a11) + (a16 - a15));
    const float gyz = 2.0f * (a03 + a04) + ((a14 - a12) + (a15 + a16));
    const float hyz = 2.0f * (a03 - a04) + ((a14 - a12) - (a15 + a16));
    const float pxy = exy2 + 6.0f * a06, mxy = exy2 - 6.0f * a06;
    const float pxz = exz2 + 6.0f * a08, mxz = exz2 - 6.0f * a08;
    const float pyz = eyz2 + 6.0f * a09, myz = eyz2 - 6.0f * a09;
    p[Q07] = pxy + gxy;
    p[Q08] = pxy - gxy;
    p[Q13] = mxy + hxy;
    p[Q14] = mxy - hxy;
    p[Q09] = pxz + gxz;
    p[Q10] = pxz - gxz;
    p[Q15] = mxz + hxz;
    p[Q16] = mxz - hxz;
    p[Q11] = pyz + gyz;
    p[Q12] = pyz - gyz;
    p[Q17] = myz + hyz;
    p[Q18] = myz - hyz;
}

// Smoothed-delta weight per axis (tdlbcpp/src/Forcing.hpp:18-35), evaluated in double.
template <typename T>
TDLB_HD inline T ib_weight(T xin) {
    T w = T(0.);
    const double x = (double)(xin < T(0) ? -xin : xin);
    if (x <= 1.5) {
        if (x <= 0.5) w = (T)((1.0 / 3.0) * (1.0 + sqrt(-3.0 * x * x + 1.0)));
        else w = (T)((1.0 / 6.0) * (5.0 - (3.0 * x) - sqrt((-3.0 * ((1.0 - x) * (1.0 - x))) + 1.0)));
    }
    return w;
}

// Reduced-precision population storage: a population is kept as its deviation
// from the rest population w_d*rho0 (w = 1/12 on the 6 faces, 1/24 on the 12
// edges, the rest populations of this lattice), scaled per direction by
// k_d = 2 on the faces and 4 on the edges: s = k_d f - c0 with c0 = rho0/6
// (k_d w_d = 1/6 in both), O(density and velocity fluctuations), inside the
// normal range of fp16 and bf16.  k_d is a power of two and c0 sits on the
// float grid of k_d f (both near 1/6), so decoding (s + c0) / k_d and encoding
// k_d f - c0 are exact in float for every stored value near rest: the 16-bit
// rounding is the only one.  (A 1/w_d scale multiplied every population by
// 12 fl(1/12) = 1 + 3e-8 per step; scales 16/32 put fp16's finest steps below
// the float grid of f, where the decode's ties round away from rest.)
// Arithmetic stays in FP32.
TDLB_HD constexpr float shift_scale(int d) { return d < 6 ? 2.0f : 4.0f; }
TDLB_HD inline float shift_offset(float rho0) { return rho0 * (1.0f / 6.0f); }

}  // namespace tdlb
