// Device-side MHD arithmetic of the HIP time step.
//
// Every expression follows the reference's evaluation order (left-to-right
// sums and products, the same divisions), so that the library built with
// -ffp-contract=off (HOWMHD_EXACT=1) rounds each value exactly as the
// reference does.  The product build (HOWMHD_EXACT=0) lets the compiler fuse
// multiply-adds and replaces divisions by constants and the WENO weight
// normalisation with reciprocal forms.
//
// Reference: src/ssprk.cpp 458-516 (primitives), 538-595 (cell eigenvalues and
// fluxes), 602-818 (interface eigenvectors), 841-945 (LLF-WENO5 flux),
// 967-994 (CT flux terms).
#pragma once

#include <hip/hip_runtime.h>

#ifndef HOWMHD_EXACT
#define HOWMHD_EXACT 0
#endif

namespace howmhd {

template <typename R>
__device__ __forceinline__ R sign1(R x) { return (x >= R(0)) ? R(1) : R(-1); }

template <typename R>
__device__ __forceinline__ R rmax(R a, R b) { return fmax(a, b); }
template <>
__device__ __forceinline__ float rmax<float>(float a, float b) { return fmaxf(a, b); }

// Square root: FP64 the IEEE square root; FP32 (the FP32 form only) the
// hardware v_sqrt_f32, within 1 ulp, instead of the correctly rounded
// sequence (range scaling, a +-1 ulp correction and a class select, 15
// instructions).
template <typename R>
__device__ __forceinline__ R rsqrt_(R x) { return sqrt(x); }
template <>
__device__ __forceinline__ float rsqrt_<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename R>
__device__ __forceinline__ R rabs(R x) { return fabs(x); }
template <>
__device__ __forceinline__ float rabs<float>(float x) { return fabsf(x); }

// a / b from y = RN(1/b), correctly rounded: q = a*y is within one ulp of
// a/b, the FMA remainder r = a - b q is exact, and RN(q + r y) = RN(a/b)
// (Markstein's theorem; barring overflow and underflow).  When r == 0 the
// quotient q is exact (and keeps the sign of a zero numerator).  So the
// result is the IEEE quotient bit for bit, in 6 instructions instead of an
// 11-instruction division when y is shared or constant.
template <typename R>
__device__ __forceinline__ R div_by(R a, R b, R y) {
    const R q = a * y;
    const R r = __builtin_fma(-b, q, a);
    return (r == R(0)) ? q : __builtin_fma(r, y, q);
}

// IEEE division a / b in parts, for operands whose division needs no range
// scaling.  The correctly rounded FP64 division sequence of gfx9 (v_div_scale,
// v_rcp_f64, two Newton steps, q = a r, remainder fma(-b, q, a), v_div_fmas,
// v_div_fixup) leaves its operands unscaled and the quotient untouched when
// b and a/b are normal, 1/b is normal, |a| >= 2^-969 and exponent(a) -
// exponent(b) < 768; the arithmetic is then div_recip (the refined
// reciprocal, a function of b alone) and div_with.  So quotients that share a
// denominator share its refinement and stay bit-for-bit the IEEE quotient.
__device__ __forceinline__ double div_recip(double b) {
    const double r0 = __builtin_amdgcn_rcp(b);
    const double r1 = __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
    return __builtin_fma(r1, __builtin_fma(-b, r1, 1.0), r1);
}
__device__ __forceinline__ double div_with(double a, double b, double r) {
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}


// FP64 1/x from v_rcp_f64 and two Newton-Raphson steps, within an ulp of the
// quotient for normal x (no range scaling; x = 0 gives NaN, not inf).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/x in the FMA forms: FP32 takes the hardware reciprocal (v_rcp_f32,
// within 1 ulp) instead of the IEEE division sequence; FP64 divides.
template <typename R>
__device__ __forceinline__ R rcp_fast(R x) { return R(1) / x; }
template <>
__device__ __forceinline__ float rcp_fast<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Division form of the reference-rounding WENO weights: IEEE divisions, or
// div_recip/div_with, which hold for t_k = (eps + IS_k)^2 <= 1e200 (then
// alpha_k in [1e-200, 6e20] and omega in (5e-222, 1]).  The unscaled form
// collects in `bad` (a wavefront mask, kept in scalar registers) the lanes
// where the high word of a t_k bit pattern reaches kUnscaledHi, the high word
// of 1e200 (every double with a smaller high word is below 1e200; t_k > 0,
// and inf and NaN compare above every finite value); the caller recomputes
// those with IEEE divisions.
enum DivMode { DIV_IEEE = 0, DIV_UNSCALED = 1 };
constexpr uint32_t kUnscaledHi = 0x6974E718u;
__device__ __forceinline__ uint32_t hi_word(double x) { return (uint32_t)((unsigned long long)__double_as_longlong(x) >> 32); }

// a / b for b > 0 from y = RN(1/b) > 0: div_by without its select.  With
// nr = fma(b, q, -a) = -r exactly, fma(-nr, y, q) equals div_by's
// (r == 0) ? q : fma(r, y, q) bit for bit whenever y > 0, signed zeros
// included (checked on 4e8 pairs: random b > 0 and a, zeros, exact multiples,
// all-ones significands); for y < 0 the two differ in the sign of a zero
// quotient, so callers check y > 0.
template <typename R>
__device__ __forceinline__ R div_pos(R a, R b, R y) {
    const R q = a * y;
    return __builtin_fma(-__builtin_fma(b, q, -a), y, q);
}

// Reference-rounding quotient by a shared reciprocal in division mode DM:
// div_pos in the unscaled mode (the caller raises its flag when y <= 0), else
// div_by.
template <int DM, typename R>
__device__ __forceinline__ R div_q(R a, R b, R y) {
    if constexpr (DM == DIV_UNSCALED) return div_pos<R>(a, b, y);
    else return div_by<R>(a, b, y);
}

// FP64 square root without its range scaling, for the reference-rounding
// sweep proper.  gfx9's correctly rounded square root scales an operand below
// 2^-767 by 2^256 (and the result by 2^-128), refines v_rsq_f64 with the
// steps below, and returns the operand itself for +-0 and +inf.  For every
// other operand >= 2^-767, and for NaN, the scaling is the identity, so these
// steps give its bits; an operand in (0, 2^-767) or negative raises `bad`,
// and the caller recomputes the tile with sqrt().
__device__ __forceinline__ double sqrt_unscaled(double x, bool& bad) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r;
    double h = r * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    const double d0 = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, e, h);
    g = __builtin_fma(d0, h, g);
    const double d1 = __builtin_fma(-g, g, x);
    g = __builtin_fma(d1, h, g);
    const bool special = __builtin_amdgcn_class(x, 0x260);  // -0, +0, +inf
    bad = bad || (x < 0x1p-767 && !special);
    return special ? x : g;
}

// Square root of the reference-rounding sweep in division mode DM: the
// unscaled form (raising `bad` out of its range) in the sweep proper, else
// the IEEE square root.
template <int DM>
__device__ __forceinline__ double sqrt_q(double x, bool& bad) {
    if constexpr (DM == DIV_UNSCALED) return sqrt_unscaled(x, bad);
    else return sqrt(x);
}

// x / c for a compile-time constant c > 0 with inv_c = RN(1/c).  Reference
// rounding: the correctly rounded quotient of div_by without its select:
// with nr = fma(c, q, -x) = -r exactly, fma(-nr, inv_c, q) equals div_by's
// (r == 0) ? q : fma(r, inv_c, q) bit for bit when inv_c > 0, signed zeros
// included (checked on 4e8 pairs for the divisors used, zeros, subnormals and
// exact multiples among them).  Like div_by it is exact barring underflow
// (subnormal x).  FMA form: multiplication by the rounded reciprocal (differs
// by at most one rounding).
template <typename R>
__device__ __forceinline__ R div_c(R x, R c, R inv_c) {
#if HOWMHD_EXACT
    const R q = x * inv_c;
    const R nr = __builtin_fma(c, q, -x);
    return __builtin_fma(-nr, inv_c, q);
#else
    (void)c;
    return x * inv_c;
#endif
}

// ----------------------------------------------------------------- cell data
// Per-cell values a sweep needs, in sweep order (n = normal, t1/t2 =
// transverse): the conserved values after the primitive floors, the physical
// flux, the wave-speed magnitudes and the primitives the interface averages
// use.  Field indices of the LDS tile:
enum CellField {
    C_RHO = 0, C_MN, C_MT1, C_MT2, C_BT1, C_BT2, C_E,      // U (projection input)
    C_F0, C_F1, C_F2, C_F3, C_F4, C_F5, C_F6,              // F
    C_VN, C_LF, C_LA, C_LS,                                // eigenvalue inputs
    C_VT1, C_VT2, C_BN, C_P, C_H,                          // interface averages
    C_NFIELD
};

// primit_x_all + eigenst_x_all centre for one cell (src/ssprk.cpp:470-593).
// Input: conserved in sweep order, FP64.  The primitive conversion, the
// floors and E -> p stay in FP64; the outputs are rounded to R.
// DM: division mode of the reference-rounding form; in the unscaled mode
// *bad is raised for a cell whose shared reciprocals are not positive or
// whose square-root operands leave the range of sqrt_unscaled (it is then
// recomputed with div_by and sqrt()).
template <typename R, int DM = DIV_IEEE>
__device__ __forceinline__ void cell_values(const double* qs, double gam, double rhomin, double pgmin,
                                            R* out, bool* bad = nullptr) {
    const double DD = qs[0], Mx = qs[1], My = qs[2], Mz = qs[3];
    const double Bx = qs[4], By = qs[5], Bz = qs[6];
    double EE = qs[7];
    double rho = DD;
#if HOWMHD_EXACT
    // correctly rounded quotients from shared reciprocals (div_by)
    const double iDD = 1.0 / DD;
    if constexpr (DM == DIV_UNSCALED) *bad = *bad || !(iDD > 0.0);  // rho >= DD > 0 then
    const double vx = div_q<DM>(Mx, DD, iDD);
    const double vy = div_q<DM>(My, DD, iDD);
    const double vz = div_q<DM>(Mz, DD, iDD);
#else
    // FP32 form: 1/rho from the hardware reciprocal and two Newton steps
    // (within an ulp of FP64; the quantities leave this function in FP32)
    double irho = (sizeof(R) == 4) ? rcp_nr(rho) : 1.0 / rho;
    const double vx = Mx * irho;
    const double vy = My * irho;
    const double vz = Mz * irho;
#endif
    const double vv2 = vx * vx + vy * vy + vz * vz;
    const double BB2 = Bx * Bx + By * By + Bz * Bz;
    double pg = (gam - 1.0) * (EE - 0.5 * (rho * vv2 + BB2));
    double Dq = DD;
    if (rho < rhomin || pg < pgmin) {
        rho = fmax(rhomin, rho);
        pg = fmax(pgmin, pg);
        EE = pg / (gam - 1.0) + 0.5 * (rho * vv2 + BB2);
        Dq = rho;
#if !HOWMHD_EXACT
        irho = (sizeof(R) == 4) ? rcp_nr(rho) : 1.0 / rho;
#endif
    }
#if HOWMHD_EXACT
    const double irho = (rho == DD) ? iDD : 1.0 / rho;  // floored rho
    const double HH = div_q<DM>(EE + pg, rho, irho);
#else
    const double HH = (EE + pg) * irho;
#endif

    // eigenst centre, in R
    const R rrho = R(rho), rvx = R(vx), rvy = R(vy), rvz = R(vz);
    const R rBx = R(Bx), rBy = R(By), rBz = R(Bz), rpg = R(pg);
    const R rDD = R(Dq), rMx = R(Mx), rMy = R(My), rMz = R(Mz), rEE = R(EE);
    const R rgam = R(gam);
    const R rBB2 = rBx * rBx + rBy * rBy + rBz * rBz;
    const R pt = rpg + R(0.5) * rBB2;
    const R vdB = rvx * rBx + rvy * rBy + rvz * rBz;
#if HOWMHD_EXACT
    static_assert(sizeof(R) == 8, "the reference-rounding form is FP64");
    const R bbn2 = div_q<DM>(rBB2, rrho, irho);
    const R bnx2 = div_q<DM>(rBx * rBx, rrho, irho);
    const R cs2 = rmax(R(0), rgam * rabs(div_q<DM>(rpg, rrho, irho)));
    // root and lf take the unscaled square root; la and ls keep the IEEE one:
    // their operands scale with B_n^2, which a blast front's precursor takes
    // below 2^-767 on whole tiles (they would all be recomputed)
    bool sq_bad = false;
    const R root = sqrt_q<DM>(rmax(R(0), (bbn2 + cs2) * (bbn2 + cs2) - R(4) * bnx2 * cs2), sq_bad);
    const R lf = sqrt_q<DM>(rmax(R(0), R(0.5) * (bbn2 + cs2 + root)), sq_bad);
    const R la = rsqrt_(rmax(R(0), bnx2));
    const R ls = rsqrt_(rmax(R(0), R(0.5) * (bbn2 + cs2 - root)));
    if constexpr (DM == DIV_UNSCALED) *bad = *bad || sq_bad;
#else
    R lf, la, ls;
    if constexpr (sizeof(R) == 8) {
        const R rirho = R(irho);
        const R bbn2 = rBB2 * rirho;
        const R bnx2 = rBx * rBx * rirho;
        const R cs2 = rmax(R(0), rgam * rabs(rpg * rirho));
        const R root = rsqrt_(rmax(R(0), (bbn2 + cs2) * (bbn2 + cs2) - R(4) * bnx2 * cs2));
        lf = rsqrt_(rmax(R(0), R(0.5) * (bbn2 + cs2 + root)));
        la = rsqrt_(rmax(R(0), bnx2));
        ls = rsqrt_(rmax(R(0), R(0.5) * (bbn2 + cs2 - root)));
    } else {
        // FP32: the same speeds from cancellation-free forms,
        // root^2 = (b^2 - cs^2)^2 + 4 bt^2 cs^2 and cs_slow^2 = bn^2 cs^2 / cf^2
        const R rirho = R(irho);
        const R bnx2 = rBx * rBx * rirho;
        const R btr2 = (rBy * rBy + rBz * rBz) * rirho;
        const R bbn2 = bnx2 + btr2;
        const R cs2 = rmax(R(0), rgam * rabs(rpg * rirho));
        const R dd = bbn2 - cs2;
        const R root = rsqrt_(dd * dd + R(4) * btr2 * cs2);
        const R lf2 = R(0.5) * (bbn2 + cs2 + root);
        lf = rsqrt_(lf2);
        la = rsqrt_(bnx2);
        ls = (lf2 > R(0)) ? rsqrt_(bnx2 * cs2 * rcp_fast(lf2)) : R(0);
    }
#endif

    out[C_RHO] = rDD;  out[C_MN] = rMx;  out[C_MT1] = rMy;  out[C_MT2] = rMz;
    out[C_BT1] = rBy;  out[C_BT2] = rBz; out[C_E] = rEE;
    out[C_F0] = rDD * rvx;
    out[C_F1] = rMx * rvx + pt - rBx * rBx;
    out[C_F2] = rMy * rvx - rBx * rBy;
    out[C_F3] = rMz * rvx - rBx * rBz;
    out[C_F4] = rBy * rvx - rBx * rvy;
    out[C_F5] = rBz * rvx - rBx * rvz;
    out[C_F6] = (rEE + pt) * rvx - rBx * vdB;
    out[C_VN] = rvx;  out[C_LF] = lf;  out[C_LA] = la;  out[C_LS] = ls;
    out[C_VT1] = rvy; out[C_VT2] = rvz; out[C_BN] = rBx; out[C_P] = rpg; out[C_H] = R(HH);
    (void)rrho;
}

// ------------------------------------------------------ interface eigensystem
// Constants of the equation of state the eigenvectors use, computed once on
// the host in FP64 with the reference's expressions: gam2 = (gam-2)/(gam-1)
// (src/ssprk.cpp:729-787), gm1 = gam-1 and inv_gm1 = RN(1/gm1).  The IEEE
// quotient is the same on the host and the device, so gam2 has the bits the
// reference computes per interface.
struct GasConst {
    double gam2, gm1, inv_gm1;
};

// Scalars of the interface eigensystem (src/ssprk.cpp:610-659).  Rows of L
// and columns of R are built from them one wave at a time.
//
// The sign flip of src/ssprk.cpp:789-818 (rows 2, 4 of L and columns 2, 4 of
// R times sign(B_t) when c_s >= c_a, else rows and columns 0, 6) is not
// applied: it cancels exactly.  Negating a row of L negates the wave's six
// projected values exactly (rounding to nearest is symmetric), the WENO flux
// of the wave is odd in them (the smoothness indicators square the
// differences), so its flux F_s is negated exactly, and the flipped column
// times the flipped flux is the unflipped product.  The only difference is
// the sign of a zero F_s, and adding a zero product to a dF component leaves
// it unchanged, because dF starts at +0 and a sum of doubles is -0 only when
// both addends are.
template <typename R>
struct Eig {
    R vx, vy, vz, By, Bz, vv2, cs, cs2, lf, ls, af, as, bty, btz, sgnBx, sqrtrho, invsqrtrho;
    R nrm_half, nrm_ent;  // FMA form: 0.5/cs2 and (gam-1)/cs2; reference rounding: nrm_half = RN(1/cs2)
};

// FP32 form: fast/slow speeds and the alpha_f, alpha_s normalisations of
// src/ssprk.cpp:627-656 from cancellation-free forms of the same quantities
// (b^2 = bn^2 + bt^2 per unit density):
//   root^2 = (b^2 - cs^2)^2 + 4 bt^2 cs^2,  cf^2 = (b^2 + cs^2 + root)/2,
//   cs_slow^2 = bn^2 cs^2 / cf^2,
//   af^2 = (cs^2 - b^2 + root) / (2 root),  as^2 = (b^2 - cs^2 + root) / (2 root),
// where a numerator that would cancel is taken as 4 bt^2 cs^2 / (root + |b^2 - cs^2|).
// In FP32 the reference forms lose every digit near the degenerate points and
// give eigenvectors with L R far from I.
template <typename R>
__device__ __forceinline__ void eig_scalars_fp32_speeds(Eig<R>& e, R cs2, R bn2, R bt2, R b2) {
    const R c2 = rmax(R(0), cs2);
    const R dd = b2 - c2;
    const R root = rsqrt_(dd * dd + R(4) * bt2 * c2);
    const R lf2 = R(0.5) * (b2 + c2 + root);
    e.lf = rsqrt_(lf2);
    e.ls = (lf2 > R(0)) ? rsqrt_(bn2 * c2 * rcp_fast(lf2)) : R(0);
    if (root >= R(1.0e-30)) {
        const R q = R(4) * bt2 * c2;
        const R fnum = (dd <= R(0)) ? root - dd : q * rcp_fast(root + dd);   // cs^2 - b^2 + root
        const R snum = (dd >= R(0)) ? root + dd : q * rcp_fast(root - dd);   // b^2 - cs^2 + root
        const R inv = R(0.5) * rcp_fast(root);
        e.af = rsqrt_(rmax(R(0), fnum * inv));
        e.as = rsqrt_(rmax(R(0), snum * inv));
    } else {
        e.af = R(1);
        e.as = R(1);
    }
}

// DM: division mode of the reference-rounding form; in the unscaled mode the
// square roots take sqrt_unscaled and raise *sqbad out of its range.
template <typename R, int DM = DIV_IEEE>
__device__ __forceinline__ void eig_scalars(R gam, const R* u0, const R* u1, Eig<R>& e, bool* sqbad = nullptr) {
#if HOWMHD_EXACT
    bool sb = false;
    auto sq = [&](R x) -> R { return sqrt_q<DM>(x, sb); };
#else
    auto sq = [](R x) -> R { return rsqrt_(x); };
#endif
    // u = (rho, vn, vt1, vt2, Bn, Bt1, Bt2, p, H) of the two cells
    const R rho = R(0.5) * (u0[0] + u1[0]);
    const R vx = R(0.5) * (u0[1] + u1[1]);
    const R vy = R(0.5) * (u0[2] + u1[2]);
    const R vz = R(0.5) * (u0[3] + u1[3]);
    const R Bx = R(0.5) * (u0[4] + u1[4]);
    const R By = R(0.5) * (u0[5] + u1[5]);
    const R Bz = R(0.5) * (u0[6] + u1[6]);
    const R HH = R(0.5) * (u0[8] + u1[8]);
    const R vv2 = vx * vx + vy * vy + vz * vz;
    const R BB2 = Bx * Bx + By * By + Bz * Bz;
#if HOWMHD_EXACT
    const R bbn2 = BB2 / rho;
    const R bnx2 = Bx * Bx / rho;
#else
    const R irho = rcp_fast(rho);
    const R bbn2 = BB2 * irho;
    const R bnx2 = Bx * Bx * irho;
#endif
    const R cs2 = (gam - R(1)) * (HH - R(0.5) * (vv2 + bbn2));
    const R cs = sq(rmax(R(0), cs2));
    const R Bt2 = By * By + Bz * Bz;
#if !HOWMHD_EXACT
    if constexpr (sizeof(R) == 4) {
        eig_scalars_fp32_speeds(e, cs2, bnx2, Bt2 * irho, bbn2);
    } else
#endif
    {
    // unscaled square roots (sq) only where the operand stays far above
    // 2^-767 in MHD flows: c_s^2, root^2, c_f^2, c_f^2 - c_slow^2 (>= 1e-30
    // here), B_t^2 (>= 1e-30) and rho.  c_a^2 and c_slow^2 scale with B_n^2,
    // and the alpha numerators vanish where B_t does: IEEE square roots.
    const R root = sq(rmax(R(0), (bbn2 + cs2) * (bbn2 + cs2) - R(4) * bnx2 * cs2));
    const R lf = sq(rmax(R(0), R(0.5) * (bbn2 + cs2 + root)));
    const R ls = rsqrt_(rmax(R(0), R(0.5) * (bbn2 + cs2 - root)));
    e.lf = lf; e.ls = ls;
    R af, as;
    if ((lf * lf - ls * ls) >= R(1.0e-30)) {
#if HOWMHD_EXACT
        const R sd = sq(lf * lf - ls * ls);
        af = rsqrt_(rmax(R(0), cs2 - ls * ls)) / sd;
        as = rsqrt_(rmax(R(0), lf * lf - cs2)) / sd;
#else
        const R rd = rcp_fast(rsqrt_(lf * lf - ls * ls));
        af = rsqrt_(rmax(R(0), cs2 - ls * ls)) * rd;
        as = rsqrt_(rmax(R(0), lf * lf - cs2)) * rd;
#endif
    } else {
        af = R(1);
        as = R(1);
    }
    e.af = af; e.as = as;
    }
    R bty, btz;
    if (Bt2 >= R(1.0e-30)) {
#if HOWMHD_EXACT
        const R sbt = sq(Bt2);
        bty = By / sbt;
        btz = Bz / sbt;
#else
        const R rb = rcp_fast(rsqrt_(Bt2));
        bty = By * rb;
        btz = Bz * rb;
#endif
    } else {
        bty = R(1) / rsqrt_(R(2));
        btz = R(1) / rsqrt_(R(2));
    }
    const R sqrtrho = sq(rho);
    e.vx = vx; e.vy = vy; e.vz = vz; e.By = By; e.Bz = Bz; e.vv2 = vv2;
    e.cs = cs; e.cs2 = cs2;
    e.bty = bty; e.btz = btz; e.sgnBx = sign1(Bx);
    e.sqrtrho = sqrtrho;
#if HOWMHD_EXACT
    e.invsqrtrho = R(1) / sqrtrho;
#else
    e.invsqrtrho = rcp_fast(sqrtrho);
#endif
#if !HOWMHD_EXACT
    const R ics2 = rcp_fast(cs2);
    e.nrm_half = R(0.5) * ics2;
    e.nrm_ent = (gam - R(1)) * ics2;
#else
    e.nrm_half = R(1) / cs2;   // RN(1/cs2): the L-row normalisation divides by cs2 through div_by
    e.nrm_ent = R(0);
    if (sqbad) *sqbad = *sqbad || sb;
#endif
}

// Row m of L after normalisation (src/ssprk.cpp:662-726).
// DM: division mode of the reference-rounding normalisation (div_q; the
// unscaled mode needs e.nrm_half > 0, which the caller checks).
template <int M, typename R, int DM = DIV_IEEE>
__device__ __forceinline__ void l_row(R gam, const Eig<R>& e, R* L, const GasConst& gc) {
    const R gam0 = R(1) - gam;
    const R gam1 = R(0.5) * (gam - R(1));
    const R af = e.af, as = e.as, lf = e.lf, ls = e.ls, cs = e.cs, cs2 = e.cs2;
    const R vx = e.vx, vy = e.vy, vz = e.vz, By = e.By, Bz = e.Bz, vv2 = e.vv2;
    const R bty = e.bty, btz = e.btz, sgnBx = e.sgnBx, sqrtrho = e.sqrtrho;
    if constexpr (M == 0) {
        L[0] = af * (gam1 * vv2 + lf * vx) - as * ls * (bty * vy + btz * vz) * sgnBx;
        L[1] = af * (gam0 * vx - lf);
        L[2] = gam0 * af * vy + as * ls * bty * sgnBx;
        L[3] = gam0 * af * vz + as * ls * btz * sgnBx;
        L[4] = gam0 * af * By + cs * as * bty * sqrtrho;
        L[5] = gam0 * af * Bz + cs * as * btz * sqrtrho;
        L[6] = -gam0 * af;
    } else if constexpr (M == 1) {
        L[0] = btz * vy - bty * vz;
        L[1] = R(0);
        L[2] = -btz;
        L[3] = bty;
        L[4] = -btz * sgnBx * sqrtrho;
        L[5] = bty * sgnBx * sqrtrho;
        L[6] = R(0);
    } else if constexpr (M == 2) {
        L[0] = as * (gam1 * vv2 + ls * vx) + af * lf * (bty * vy + btz * vz) * sgnBx;
        L[1] = gam0 * as * vx - as * ls;
        L[2] = gam0 * as * vy - af * lf * bty * sgnBx;
        L[3] = gam0 * as * vz - af * lf * btz * sgnBx;
        L[4] = gam0 * as * By - cs * af * bty * sqrtrho;
        L[5] = gam0 * as * Bz - cs * af * btz * sqrtrho;
        L[6] = -gam0 * as;
    } else if constexpr (M == 3) {
#if HOWMHD_EXACT
        // -cs2/gam0 = cs2/(gam-1) exactly (RN(1-gam) = -RN(gam-1)); the
        // sweep proper divides by the constant through its reciprocal
        // (div_pos: correctly rounded, barring a subnormal cs2)
        if constexpr (DM == DIV_UNSCALED) L[0] = div_pos<R>(cs2, R(gc.gm1), R(gc.inv_gm1)) - R(0.5) * vv2;
        else L[0] = -cs2 / gam0 - R(0.5) * vv2;
#else
        L[0] = -cs2 * rcp_fast(gam0) - R(0.5) * vv2;
#endif
        L[1] = vx;
        L[2] = vy;
        L[3] = vz;
        L[4] = By;
        L[5] = Bz;
        L[6] = R(-1);
    } else if constexpr (M == 4) {
        L[0] = as * (gam1 * vv2 - ls * vx) - af * lf * (bty * vy + btz * vz) * sgnBx;
        L[1] = as * (gam0 * vx + ls);
        L[2] = gam0 * as * vy + af * lf * bty * sgnBx;
        L[3] = gam0 * as * vz + af * lf * btz * sgnBx;
        L[4] = gam0 * as * By - cs * af * bty * sqrtrho;
        L[5] = gam0 * as * Bz - cs * af * btz * sqrtrho;
        L[6] = -gam0 * as;
    } else if constexpr (M == 5) {
        L[0] = btz * vy - bty * vz;
        L[1] = R(0);
        L[2] = -btz;
        L[3] = bty;
        L[4] = btz * sgnBx * sqrtrho;
        L[5] = -bty * sgnBx * sqrtrho;
        L[6] = R(0);
    } else {
        L[0] = af * (gam1 * vv2 - lf * vx) + as * ls * (bty * vy + btz * vz) * sgnBx;
        L[1] = af * (gam0 * vx + lf);
        L[2] = gam0 * af * vy - as * ls * bty * sgnBx;
        L[3] = gam0 * af * vz - as * ls * btz * sgnBx;
        L[4] = gam0 * af * By + cs * as * bty * sqrtrho;
        L[5] = gam0 * af * Bz + cs * as * btz * sqrtrho;
        L[6] = -gam0 * af;
    }
    // normalisation, src/ssprk.cpp:718-726
#pragma unroll
    for (int c = 0; c < 7; ++c) {
#if HOWMHD_EXACT
        if constexpr (M == 1 || M == 5) L[c] = R(0.5) * L[c];
        else if constexpr (M == 3) L[c] = div_q<DM>(-gam0 * L[c], cs2, e.nrm_half);
        else L[c] = div_q<DM>(R(0.5) * L[c], cs2, e.nrm_half);
#else
        if constexpr (M == 1 || M == 5) L[c] = R(0.5) * L[c];
        else if constexpr (M == 3) L[c] = L[c] * e.nrm_ent;
        else L[c] = L[c] * e.nrm_half;
#endif
    }
    // (the sign flip of src/ssprk.cpp:796-806 cancels against that of R; see Eig)
}

// Column m of R (src/ssprk.cpp:729-787).
template <int M, typename R>
__device__ __forceinline__ void r_col(R gam, const Eig<R>& e, R* Rc, const GasConst& gc) {
#if HOWMHD_EXACT
    (void)gam;
    const R gam2 = R(gc.gam2);   // (gam - 2)/(gam - 1), computed on the host
#else
    (void)gc;
    const R gam2 = (gam - R(2)) * rcp_fast(gam - R(1));
#endif
    const R af = e.af, as = e.as, lf = e.lf, ls = e.ls, cs = e.cs, cs2 = e.cs2;
    const R vx = e.vx, vy = e.vy, vz = e.vz, vv2 = e.vv2;
    const R bty = e.bty, btz = e.btz, sgnBx = e.sgnBx, isr = e.invsqrtrho;
    if constexpr (M == 0) {
        Rc[0] = af;
        Rc[1] = af * (vx - lf);
        Rc[2] = af * vy + as * ls * bty * sgnBx;
        Rc[3] = af * vz + as * ls * btz * sgnBx;
        Rc[4] = cs * as * bty * isr;
        Rc[5] = cs * as * btz * isr;
        Rc[6] = af * (lf * lf - lf * vx + R(0.5) * vv2 - gam2 * cs2) + as * ls * (bty * vy + btz * vz) * sgnBx;
    } else if constexpr (M == 1) {
        Rc[0] = R(0);
        Rc[1] = R(0);
        Rc[2] = -btz;
        Rc[3] = bty;
        Rc[4] = -btz * sgnBx * isr;
        Rc[5] = bty * sgnBx * isr;
        Rc[6] = bty * vz - btz * vy;
    } else if constexpr (M == 2) {
        Rc[0] = as;
        Rc[1] = as * (vx - ls);
        Rc[2] = as * vy - af * lf * bty * sgnBx;
        Rc[3] = as * vz - af * lf * btz * sgnBx;
        Rc[4] = -cs * af * bty * isr;
        Rc[5] = -cs * af * btz * isr;
        Rc[6] = as * (ls * ls - ls * vx + R(0.5) * vv2 - gam2 * cs2) - af * lf * (bty * vy + btz * vz) * sgnBx;
    } else if constexpr (M == 3) {
        Rc[0] = R(1);
        Rc[1] = vx;
        Rc[2] = vy;
        Rc[3] = vz;
        Rc[4] = R(0);
        Rc[5] = R(0);
        Rc[6] = R(0.5) * vv2;
    } else if constexpr (M == 4) {
        Rc[0] = as;
        Rc[1] = as * (vx + ls);
        Rc[2] = as * vy + af * lf * bty * sgnBx;
        Rc[3] = as * vz + af * lf * btz * sgnBx;
        Rc[4] = -cs * af * bty * isr;
        Rc[5] = -cs * af * btz * isr;
        Rc[6] = as * (ls * ls + ls * vx + R(0.5) * vv2 - gam2 * cs2) + af * lf * (bty * vy + btz * vz) * sgnBx;
    } else if constexpr (M == 5) {
        Rc[0] = R(0);
        Rc[1] = R(0);
        Rc[2] = -btz;
        Rc[3] = bty;
        Rc[4] = btz * sgnBx * isr;
        Rc[5] = -bty * sgnBx * isr;
        Rc[6] = bty * vz - btz * vy;
    } else {
        Rc[0] = af;
        Rc[1] = af * (vx + lf);
        Rc[2] = af * vy - as * ls * bty * sgnBx;
        Rc[3] = af * vz - as * ls * btz * sgnBx;
        Rc[4] = cs * as * bty * isr;
        Rc[5] = cs * as * btz * isr;
        Rc[6] = af * (lf * lf + lf * vx + R(0.5) * vv2 - gam2 * cs2) - as * ls * (bty * vy + btz * vz) * sgnBx;
    }
    // (the sign flip of src/ssprk.cpp:808-818 cancels against that of L; see Eig)
}

// ------------------------------------------------------------------ WENO5
// One side of the Jiang-Shu WENO5 correction (src/ssprk.cpp:891-912):
// phi_N(a,b,c,d) = w0 (a-2b+c)/3 + (w2-1/2)(b-2c+d)/6 with
// IS0 = 13(a-b)^2 + 3(a-3b)^2, IS1 = 13(b-c)^2 + 3(b+c)^2,
// IS2 = 13(c-d)^2 + 3(3c-d)^2, alpha = (1,6,3)/(eps+IS)^2, eps = 1e-10.
template <typename R, int DM = DIV_IEEE>
__device__ __forceinline__ R weno_side(R a, R b, R c, R d, unsigned long long& bad, R* w0_out = nullptr, R* w2_out = nullptr) {
    const R eps = R(1.0e-10);
    const R IS0 = R(13) * (a - b) * (a - b) + R(3) * (a - R(3) * b) * (a - R(3) * b);
    const R IS1 = R(13) * (b - c) * (b - c) + R(3) * (b + c) * (b + c);
    const R IS2 = R(13) * (c - d) * (c - d) + R(3) * (R(3) * c - d) * (R(3) * c - d);
#if HOWMHD_EXACT
    const R t0 = (eps + IS0) * (eps + IS0);
    const R t1 = (eps + IS1) * (eps + IS1);
    const R t2 = (eps + IS2) * (eps + IS2);
    R omega0, omega2;
    if constexpr (DM == DIV_UNSCALED) {
        bad |= __ballot(max(hi_word(t0), max(hi_word(t1), hi_word(t2))) >= kUnscaledHi);
        const R alpha0 = div_with(R(1), t0, div_recip(t0));
        const R alpha1 = div_with(R(6), t1, div_recip(t1));
        const R alpha2 = div_with(R(3), t2, div_recip(t2));
        const R S = alpha0 + alpha1 + alpha2;
        const R rS = div_recip(S);
        omega0 = div_with(alpha0, S, rS);
        omega2 = div_with(alpha2, S, rS);
    } else {
        const R alpha0 = R(1) / t0;
        const R alpha1 = R(6) / t1;
        const R alpha2 = R(3) / t2;
        omega0 = alpha0 / (alpha0 + alpha1 + alpha2);
        omega2 = alpha2 / (alpha0 + alpha1 + alpha2);
    }
    if (w0_out) { *w0_out = omega0; *w2_out = omega2; }
    // a - 2b is a - RN(2b) with 2b exact, so one fused multiply-add gives
    // its bits
    return div_c<R>(omega0 * (fma(R(-2), b, a) + c), R(3), R(1.0 / 3.0)) +
           div_c<R>((omega2 - R(0.5)) * (fma(R(-2), c, b) + d), R(6), R(1.0 / 6.0));
#else
    // alpha_k = d_k / t_k with t_k = (eps+IS_k)^2: the normalised weights are
    // w0 = t1 t2 / S, w2 = 3 t0 t1 / S with S = t1 t2 + 6 t0 t2 + 3 t0 t1,
    // one division instead of five.  t_k >= 1e-20 and products stay far
    // inside the double range; the float form keeps the per-alpha
    // reciprocals, whose products could leave the float range.
    const R t0 = (eps + IS0) * (eps + IS0);
    const R t1 = (eps + IS1) * (eps + IS1);
    const R t2 = (eps + IS2) * (eps + IS2);
    R omega0, omega2;
    if constexpr (sizeof(R) == 8) {
        const R t01 = t0 * t1;
        const R t12 = t1 * t2;
        const R inv = R(1) / (t12 + R(6) * t0 * t2 + R(3) * t01);
        omega0 = t12 * inv;
        omega2 = R(3) * t01 * inv;
    } else {
        const R alpha0 = rcp_fast(t0);
        const R alpha1 = R(6) * rcp_fast(t1);
        const R alpha2 = R(3) * rcp_fast(t2);
        const R inv = rcp_fast(alpha0 + alpha1 + alpha2);
        omega0 = alpha0 * inv;
        omega2 = alpha2 * inv;
    }
    if (w0_out) { *w0_out = omega0; *w2_out = omega2; }
    return omega0 * (a - R(2) * b + c) * R(1.0 / 3.0) + (omega2 - R(0.5)) * (b - R(2) * c + d) * R(1.0 / 6.0);
#endif
}

// Characteristic flux of one wave from its 6 projected values
// (src/ssprk.cpp:883-937): first - phi(plus side) + phi(minus side).
template <typename R, int DM = DIV_IEEE>
__device__ __forceinline__ R weno_wave(const R* Fsk, const R* qsk, R amx, unsigned long long& bad) {
    R dF[5], dq[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        dF[k] = Fsk[k + 1] - Fsk[k];
        dq[k] = qsk[k + 1] - qsk[k];
    }
    const R first = div_c<R>(-Fsk[1] + R(7) * Fsk[2] + R(7) * Fsk[3] - Fsk[4], R(12), R(1.0 / 12.0));
    const R second = weno_side<R, DM>(R(0.5) * (dF[0] + amx * dq[0]), R(0.5) * (dF[1] + amx * dq[1]),
                                      R(0.5) * (dF[2] + amx * dq[2]), R(0.5) * (dF[3] + amx * dq[3]), bad);
    const R third = weno_side<R, DM>(R(0.5) * (dF[4] - amx * dq[4]), R(0.5) * (dF[3] - amx * dq[3]),
                                     R(0.5) * (dF[2] - amx * dq[2]), R(0.5) * (dF[1] - amx * dq[1]), bad);
    return first - second + third;
}

// weno_wave that also raises `dev` to the largest |w_r - d_r| of both sides,
// d = (0.1, 0.6, 0.3) (the label of the learned indicator; weights of
// src/ssprk.cpp:903-908, 926-931, w1 = 1 - w0 - w2).
template <typename R, int DM = DIV_IEEE>
__device__ __forceinline__ R weno_wave_dev(const R* Fsk, const R* qsk, R amx, R& dev, unsigned long long& bad) {
    R dF[5], dq[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        dF[k] = Fsk[k + 1] - Fsk[k];
        dq[k] = qsk[k + 1] - qsk[k];
    }
    const R first = div_c<R>(-Fsk[1] + R(7) * Fsk[2] + R(7) * Fsk[3] - Fsk[4], R(12), R(1.0 / 12.0));
    R w0, w2;
    const R second = weno_side<R, DM>(R(0.5) * (dF[0] + amx * dq[0]), R(0.5) * (dF[1] + amx * dq[1]),
                                      R(0.5) * (dF[2] + amx * dq[2]), R(0.5) * (dF[3] + amx * dq[3]), bad, &w0, &w2);
    dev = rmax(dev, rmax(rabs(w0 - R(0.1)), rmax(rabs(w2 - R(0.3)), rabs(R(1) - w0 - w2 - R(0.6)))));
    const R third = weno_side<R, DM>(R(0.5) * (dF[4] - amx * dq[4]), R(0.5) * (dF[3] - amx * dq[3]),
                                     R(0.5) * (dF[2] - amx * dq[2]), R(0.5) * (dF[1] - amx * dq[1]), bad, &w0, &w2);
    dev = rmax(dev, rmax(rabs(w0 - R(0.1)), rmax(rabs(w2 - R(0.3)), rabs(R(1) - w0 - w2 - R(0.6)))));
    return first - second + third;
}

// Linear upwind-5 flux of one component with one LLF speed alpha: the WENO5
// flux of weno_wave with the ideal weights (w0 = 0.1, w2 = 0.3).  Expanded in
// the cell values it is
//   (F0 - 8 F1 + 37 F2 + 37 F3 - 8 F4 + F5 + alpha (U0 - 5 U1 + 10 U2 - 10 U3 + 5 U4 - U5)) / 60,
// a centred 6th-order flux plus the 5th-difference LLF dissipation, taken
// here in symmetric pairs.  With one alpha for all waves the characteristic
// projection cancels (R L = I), so it runs on the physical components
// directly.
template <typename R>
__device__ __forceinline__ R upwind5(const R* F, const R* U, R alpha) {
    const R sf = (F[0] + F[5]) - R(8) * (F[1] + F[4]) + R(37) * (F[2] + F[3]);
    const R su = (U[0] - U[5]) - R(5) * (U[1] - U[4]) + R(10) * (U[2] - U[3]);
    return (sf + alpha * su) * R(1.0 / 60.0);
}

}  // namespace howmhd
