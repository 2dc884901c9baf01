/*
 * vw_shared.h -- FP64 formulas evaluated bit-for-bit identically by the
 * gfx950 HIP kernels and by the CPU oracle.
 *
 * Every quantity that feeds a comparison (occlusion masks, tile-cone classes,
 * uv-grid keys, segment parameters) is computed here with a fixed sequence of
 * correctly rounded IEEE-754 operations (+ - * / sqrt fma rint floor ceil),
 * so a C compiler on the host and hipcc on gfx950 produce the same bits.
 * Translation units that include this header must be compiled with
 * -ffp-contract=off and without fast-math; fused multiply-adds are written
 * out explicitly with fma().
 *
 * Elementary functions: sin/cos are evaluated in "turns" (x means 2*pi*x
 * radians).  Reduction is exact (x - rint(x), then quarter-turn by rint(4f));
 * the reduced argument |y| <= pi/4 is fed to Taylor polynomials (sin through
 * y^17, cos through y^16; truncation < 1e-18 relative on |y| <= pi/4).  The
 * coefficients are the nearest doubles of 1/k!.  No vendor math library
 * constants are used.
 *
 * Reference semantics restated here (FP32 in the reference, FP64 here):
 *   pix2lmn_nest_kernel            viss_recon_kernel.cuh:29-112
 *   occlusion test c>=cosphi       viss_recon_kernel.cuh:270-274
 *   classify_tile_cone             mirror_tile_vissfast/viss_recon_kernel_mirror_tile_vissfast.cuh:91-98
 *   build_tile_cone_meta_kernel    mirror_tile_vissfast/viss_recon_kernel_mirror_tile_vissfast.cuh:21-81
 *   k_compute_sat1_xyz, Mt_kf      orbit_online.cuh:39-83
 *   k_compute_other_sats           orbit_online.cuh:111-161
 *   k_compute_uvw_xyz              orbit_online.cuh:186-231
 *   compute_fa_fb_host_...         viss_recon_kernel.cuh:673-701
 *   RES / du                       main_recon_online.cu:377-383
 *   build_locg_kernel (full)       viss_recon_kernel.cuh:295-312
 *   build_locg_half_kernel         recon_mirror/viss_recon_kernel_mirror.cuh:462-481
 *   phase_correct_viss_halfsym     common.hpp:255-264
 */
#ifndef VW_SHARED_H
#define VW_SHARED_H

#include <stdint.h>
#include <math.h>

#ifdef __HIPCC__
#define VW_HD __host__ __device__ static inline
#else
#define VW_HD static inline
#endif

/* ---- constants (nearest doubles, computed from exact rationals) ---- */
#define VW_TWOPI_HI 0x1.921fb54442d18p+2  /* 2*pi rounded            */
#define VW_TWOPI_LO 0x1.1a62633145c07p-52 /* 2*pi - VW_TWOPI_HI      */
#define VW_INV_TWOPI 0x1.45f306dc9c883p-3 /* 1/(2*pi) rounded        */
#define VW_PI 0x1.921fb54442d18p+1

/* tile / pair classes */
#define VW_CLS_MIXED 0
#define VW_CLS_VISIBLE 1
#define VW_CLS_HIDDEN 2

/* ------------------------------------------------------------------ */
/* elementary functions                                                */
/* ------------------------------------------------------------------ */

/* sin(2*pi*r) and cos(2*pi*r) for finite |r| < 2^51. */
VW_HD void vw_sincos_turn(double r, double* s_out, double* c_out) {
  double f = r - rint(r);          /* exact, |f| <= 1/2            */
  double q = rint(4.0 * f);        /* quadrant in {-2..2}          */
  double t = (4.0 * f - q) * 0.25; /* exact, |t| <= 1/8 turn       */
  double y = fma(t, VW_TWOPI_HI, t * VW_TWOPI_LO);
  double z = y * y;
  /* sin(y) = y + y^3 * P(z) */
  double p = 0x1.952c77030ad4ap-49;       /* +1/17! */
  p = fma(p, z, -0x1.ae7f3e733b81fp-41);  /* -1/15! */
  p = fma(p, z, 0x1.6124613a86d09p-33);   /* +1/13! */
  p = fma(p, z, -0x1.ae64567f544e4p-26);  /* -1/11! */
  p = fma(p, z, 0x1.71de3a556c734p-19);   /* +1/9!  */
  p = fma(p, z, -0x1.a01a01a01a01ap-13);  /* -1/7!  */
  p = fma(p, z, 0x1.1111111111111p-7);    /* +1/5!  */
  p = fma(p, z, -0x1.5555555555555p-3);   /* -1/3!  */
  double sy = fma(y * z, p, y);
  /* cos(y) = 1 + z * Q(z) */
  double qq = 0x1.ae7f3e733b81fp-45;      /* +1/16! */
  qq = fma(qq, z, -0x1.93974a8c07c9dp-37); /* -1/14! */
  qq = fma(qq, z, 0x1.1eed8eff8d898p-29);  /* +1/12! */
  qq = fma(qq, z, -0x1.27e4fb7789f5cp-22); /* -1/10! */
  qq = fma(qq, z, 0x1.a01a01a01a01ap-16);  /* +1/8!  */
  qq = fma(qq, z, -0x1.6c16c16c16c17p-10); /* -1/6!  */
  qq = fma(qq, z, 0x1.5555555555555p-5);   /* +1/4!  */
  qq = fma(qq, z, -0.5);                   /* -1/2!  */
  double cy = fma(z, qq, 1.0);
  /* quadrant rotation without branches: odd quadrants swap, sign flips on
   * bit 1 of q (sin) and of q+1 (cos); negation is exact */
  int qi = ((int)q) & 3;
  double a = (qi & 1) ? cy : sy;
  double b = (qi & 1) ? sy : cy;
  *s_out = (qi & 2) ? -a : a;
  *c_out = ((qi + 1) & 2) ? -b : b;
}

/* Table-driven sin/cos of 2*pi*x (Tang's method).  The argument is passed
 * as t = 1024 x (x in turns; the kernels fold the factor 1024 = 2^10 into
 * their operands, which is exact).  t = k + f with k = rint(t) and
 * |f| <= 1/2 exactly, so 2*pi*x = 2*pi*k/1024 + y with |y| <= pi/1024;
 * sin(y) = y - y^3/6 + y^5/120 (omitted term < 6e-22) and
 * cos(y) = 1 - y^2/2 + y^4/24 (omitted term < 1.2e-18), then one rotation
 * by the table entry (cos, sin)(2*pi*(k mod 1024)/1024).  The table is built
 * by vw_sincos_tab_build with vw_sincos_turn, identically on host and
 * device.  Error within ~2 ulp of sin/cos; 14 operations plus one table
 * read. */
#ifndef VW_TAB_LOG2
#define VW_TAB_LOG2 10
#endif
#define VW_TAB_N (1 << VW_TAB_LOG2)
#define VW_TAB_SCALE ((double)VW_TAB_N)
VW_HD void vw_sincos_tab_build(int k, double* tab) {
  vw_sincos_turn((double)k / (double)VW_TAB_N, &tab[2 * k + 1], &tab[2 * k]);
}

/* low 32 bits of a double's encoding */
VW_HD uint32_t vw_lo32(double d) {
  union { double d; uint64_t u; } v;
  v.d = d;
  return (uint32_t)v.u;
}

/* t -> f = t - rint(t) (exact) and the table index rint(t) mod 1024.
 * Adding 1.5*2^52 rounds t to an integer k (ties to even, as rint) held in
 * the low mantissa bits as 2^51 + k; valid for |t| < 2^51 (|x| < 2^41 turns). */
VW_HD double vw_tab_split(double t, int* idx) {
  const double M = 0x1.8p52;
  double sh = t + M;
  double k = sh - M;
  *idx = (int)(vw_lo32(sh) & (VW_TAB_N - 1));
  return t - k;
}

/* sin and cos of 2*pi*f/N, |f| <= 1/2, as polynomials in f (Taylor
 * coefficients of sin/cos(a f), a = 2*pi/N, rounded to double; the values
 * for N = 256 are scaled by exact powers of two for larger tables):
 *   sin = f (a + f^2 (-a^3/6 + f^2 a^5/120 [- f^2 a^7/5040]))
 *   cos = 1 + f^2 (-a^2/2 + f^2 (a^4/24 [- f^2 a^6/720]))
 * the bracketed terms only for smaller tables (their size < 1.2e-18 at
 * N = 1024).  Working in f saves the product y = a f. */
#if VW_TAB_LOG2 < 8 || VW_TAB_LOG2 > 10
#error "table size 2^8 .. 2^10"
#endif
#define VW_TAB_SH(k) ((double)(1ULL << ((k) * (VW_TAB_LOG2 - 8))))
VW_HD void vw_tab_poly(double f, double* sy, double* cy) {
  double z = f * f;
#if VW_TAB_LOG2 < 9
  double ps = fma(fma(z, -0x1.32d2cce62bd86p-50 / VW_TAB_SH(7), 0x1.466bc6775aae2p-34 / VW_TAB_SH(5)), z,
                  -0x1.4abbce625be53p-19 / VW_TAB_SH(3));
#else
  double ps = fma(z, 0x1.466bc6775aae2p-34 / VW_TAB_SH(5), -0x1.4abbce625be53p-19 / VW_TAB_SH(3));
#endif
  *sy = f * fma(z, ps, 0x1.921fb54442d18p-6 / VW_TAB_SH(1));
#if VW_TAB_LOG2 < 10
  double pc = fma(fma(z, -0x1.55d3c7e3cbffap-42 / VW_TAB_SH(6), 0x1.03c1f081b5ac4p-26 / VW_TAB_SH(4)), z,
                  -0x1.3bd3cc9be45dep-12 / VW_TAB_SH(2));
#else
  double pc = fma(z, 0x1.03c1f081b5ac4p-26 / VW_TAB_SH(4), -0x1.3bd3cc9be45dep-12 / VW_TAB_SH(2));
#endif
  *cy = fma(z, pc, 1.0);
}

/* The kernels' form of the same polynomials, one degree lower in each:
 *   sin = f (a - f^2 a^3/6),  cos = 1 - f^2 a^2/2   (y = a f, |y| <= pi/N)
 * at N = 1024 the omitted terms are below y^5/120 <= 2.3e-15 (sin) and
 * y^4/24 <= 3.8e-12 (cos) in absolute value per term, far inside the plain
 * form's 1e-9 bound on the outputs, for two FP64 FMAs less per term (the
 * oracle keeps the full polynomials). */
#if VW_TAB_LOG2 != 10
#error "the kernels' short polynomials are costed for a 1024-entry table"
#endif
#define VW_SHORT_S1 (0x1.921fb54442d18p-6 / VW_TAB_SH(1))   /* a */
#define VW_SHORT_S3 (-0x1.4abbce625be53p-19 / VW_TAB_SH(3)) /* -a^3/6 */
#define VW_SHORT_C2 (-0x1.3bd3cc9be45dep-12 / VW_TAB_SH(2)) /* -a^2/2 */
VW_HD void vw_tab_poly_short(double f, double* sy, double* cy) {
  const double z = f * f;
  *sy = f * fma(z, VW_SHORT_S3, VW_SHORT_S1);
  *cy = fma(z, VW_SHORT_C2, 1.0);
}

/* rotation by the table entry (ec, es) */
VW_HD void vw_tab_rotate(double ec, double es, double sy, double cy, double* s_out, double* c_out) {
  *s_out = fma(es, cy, ec * sy);
  *c_out = fma(ec, cy, -(es * sy));
}

/* sin and cos of 2*pi*t/1024 */
VW_HD void vw_sincos_tab(double t, const double* tab, double* s_out, double* c_out) {
  int i;
  double f = vw_tab_split(t, &i);
  double sy, cy;
  vw_tab_poly(f, &sy, &cy);
  vw_tab_rotate(tab[2 * i], tab[2 * i + 1], sy, cy, s_out, c_out);
}

VW_HD double vw_dot3(double ax, double ay, double az, double bx, double by, double bz) {
  return fma(az, bz, fma(ay, by, ax * bx));
}

VW_HD double vw_clamp1(double x) { return fmin(1.0, fmax(-1.0, x)); }

/* ------------------------------------------------------------------ */
/* HEALPix NESTED pixel -> direction (l, m, n)                         */
/* viss_recon_kernel.cuh:29-112: l = sin(th)cos(phi'), m = sin(th)sin(phi'), */
/* n = cos(th) = z, phi' = -phi.  z and sin(th) are formed from exact    */
/* integer ratios (one rounding for z, three for sin(th)).               */
/* ------------------------------------------------------------------ */
VW_HD void vw_pix2lmn_nest(int64_t nside, int64_t ipix, double* l, double* m, double* n) {
  int64_t npface = nside * nside;
  int64_t face = ipix / npface;
  int64_t ipf = ipix - face * npface;
  int64_t ix = 0, iy = 0;
  for (int b = 0; b < 31; ++b) {
    ix |= ((ipf >> (2 * b)) & 1) << b;
    iy |= ((ipf >> (2 * b + 1)) & 1) << b;
  }
  int64_t jrll = 2 + face / 4;
  int64_t fm = face % 4;
  int64_t jpll = (face / 4 == 1) ? 2 * fm : 2 * fm + 1;
  int64_t jrt = ix + iy;
  int64_t jpt = ix - iy;
  int64_t jr = jrll * nside - jrt - 1;
  int64_t nr, kshift;
  double z, sth;
  double d3n2 = (double)(3 * nside * nside);
  if (jr < nside) {
    nr = jr;
    kshift = 0;
    z = (double)(3 * nside * nside - nr * nr) / d3n2;
    sth = ((double)nr * sqrt((double)(6 * nside * nside - nr * nr))) / d3n2;
  } else if (jr <= 3 * nside) {
    nr = nside;
    kshift = (jr - nside) & 1;
    int64_t k = 2 * nside - jr;
    z = (double)(2 * k) / (double)(3 * nside);
    sth = sqrt((double)(9 * nside * nside - 4 * k * k)) / (double)(3 * nside);
  } else {
    nr = 4 * nside - jr;
    kshift = 0;
    z = -((double)(3 * nside * nside - nr * nr) / d3n2);
    sth = ((double)nr * sqrt((double)(6 * nside * nside - nr * nr))) / d3n2;
  }
  int64_t nl4 = 4 * nside;
  int64_t jp = (jpll * nr + jpt + 1 + kshift) >> 1;
  if (jp > nl4) jp -= nl4;
  if (jp < 1) jp += nl4;
  /* phi/(2 pi) = (2 jp - kshift - 1) / (8 nr); reduce the numerator exactly */
  int64_t Q = 2 * jp - kshift - 1;
  int64_t D = 8 * nr;
  int64_t Qr = Q % D;
  if (2 * Qr > D) Qr -= D;
  double sp, cp;
  vw_sincos_turn((double)Qr / (double)D, &sp, &cp);
  *l = sth * cp;
  *m = -(sth * sp);
  *n = z;
}

/* ------------------------------------------------------------------ */
/* occlusion                                                            */
/* ------------------------------------------------------------------ */

/* Direction (l,m,n) is visible from an antenna at P (km) iff
 * (s . P) * invn >= cosphi   (viss_recon_kernel.cuh:270-274). */
VW_HD int vw_visible(double l, double m, double n, double px, double py, double pz,
                     double invn, double cosphi) {
  double c = vw_dot3(l, m, n, px, py, pz) * invn;
  return c >= cosphi;
}

/* Antenna frame: inverse norm, unit vector and the occlusion threshold.
 * occ_radius <= 0 selects the single global threshold cosphi_global
 * (reference lunar mode, main_recon_online.cu:83-86); otherwise the per-
 * antenna cone of a sphere of radius occ_radius centred at the origin:
 * cosphi = -sqrt(1 - (R/|P|)^2). */
VW_HD void vw_antenna_frame(double px, double py, double pz, double occ_radius,
                            double cosphi_global, double* invn, double* ux, double* uy,
                            double* uz, double* cosphi) {
  double nrm = sqrt(vw_dot3(px, py, pz, px, py, pz));
  double in = 1.0 / nrm;
  *invn = in;
  *ux = px * in;
  *uy = py * in;
  *uz = pz * in;
  if (occ_radius > 0.0) {
    double q = occ_radius * in;
    *cosphi = -sqrt(fmax(0.0, fma(-q, q, 1.0)));
  } else {
    *cosphi = cosphi_global;
  }
}

/* Global lunar threshold cos(pi - asin(R/(R+h))) = -sqrt(1 - (R/(R+h))^2). */
VW_HD double vw_cosphi_from_altitude(double R, double h) {
  double q = R / (R + h);
  return -sqrt(fma(-q, q, 1.0));
}

/* One antenna against a tile cone: bit0 = all visible, bit1 = all hidden.
 * classify_tile_cone (vissfast .cuh:91-98): with d = cos(delta) the angle to
 * the cone axis and A the cone half-angle, the extremes of s.u over the cone
 * are cos(delta + A) and cos(delta - A) while delta + A <= pi and A <= delta.
 * Outside that range the cone contains -u (minimum -1) or u (maximum +1);
 * the reference formula is not conservative there (cones wider than a few
 * tens of degrees, e.g. 256-pixel tiles at Nside <= 8), so those two cases
 * clamp to -1 / +1.  For the small tiles of the production configurations
 * this is the reference's classification unchanged. */
VW_HD int vw_cone_bits(double ux, double uy, double uz, double cx, double cy, double cz,
                       double cosA, double sinA, double cosphi) {
  double d = vw_clamp1(vw_dot3(ux, uy, uz, cx, cy, cz));
  double sind = sqrt(fmax(0.0, fma(-d, d, 1.0)));
  double ss = sind * sinA;
  double lower = (cosA + d <= 0.0) ? -1.0 : fma(d, cosA, -ss);
  double upper = (cosA <= d) ? 1.0 : fma(d, cosA, ss);
  return (lower >= cosphi ? 1 : 0) | (upper < cosphi ? 2 : 0);
}

/* Pair class from the two endpoint bit sets (vissfast .cuh:521-527). */
VW_HD int vw_pair_class(int bits_a, int bits_b) {
  if ((bits_a | bits_b) & 2) return VW_CLS_HIDDEN;
  if (bits_a & bits_b & 1) return VW_CLS_VISIBLE;
  return VW_CLS_MIXED;
}

/* Tile cone from the pairwise-summed direction sum (sx,sy,sz) and the
 * minimum of c.s over the tile (computed by the caller after the centre). */
VW_HD void vw_cone_centre(double sx, double sy, double sz, double* cx, double* cy, double* cz) {
  double nrm = sqrt(vw_dot3(sx, sy, sz, sx, sy, sz));
  if (!(nrm > 0.0)) { /* directions cancel: any unit axis gives a valid cone */
    *cx = 0.0; *cy = 0.0; *cz = 1.0;
    return;
  }
  *cx = sx / nrm;
  *cy = sy / nrm;
  *cz = sz / nrm;
}

VW_HD void vw_cone_radius(double mind, double* cosA, double* sinA) {
  double ca = vw_clamp1(mind);
  *cosA = ca;
  *sinA = sqrt(fmax(0.0, fma(-ca, ca, 1.0)));
}

/* ------------------------------------------------------------------ */
/* geometry generators                                                  */
/* ------------------------------------------------------------------ */

/* Rotation of a circular orbit point (orbit_online.cuh:59-77):
 * X = r(cw cr - sw ci sr), Y = r(cw sr + sw ci cr), Z = r sw si. */
VW_HD void vw_orbit_xyz(double r, double sw, double cw, double sr, double cr, double si,
                        double ci, double* x, double* y, double* z) {
  double swci = sw * ci;
  *x = r * ((cw * cr) - (swci * sr));
  *y = r * ((cw * sr) + (swci * cr));
  *z = r * (sw * si);
}

/* Reference lunar formation (orbit_online.cuh:4-13, 249-250). */
#define VW_LUNAR_SATNUM 8
#define VW_LUNAR_A_KM 2038.14
#define VW_LUNAR_R_KM 1737.1
#define VW_LUNAR_H_KM 300.0

VW_HD double vw_lunar_r1(int s) {
  double u0 = 100e3 / 23.0;
  switch (s) {
    case 1: return u0;
    case 2: return 4.0 * u0;
    case 3: return 10.0 * u0;
    case 4: return 16.0 * u0;
    case 5: return 18.0 * u0;
    case 6: return 21.0 * u0;
    case 7: return 23.0 * u0;
    default: return 0.0;
  }
}
VW_HD double vw_lunar_r2(int s) {
  double u0 = 100e3 / 23.0;
  switch (s) {
    case 1: return 100.0;
    case 2: return 0.4 * u0;
    case 3: return 1.0 * u0;
    case 4: return 1.6 * u0;
    case 5: return 1.8 * u0;
    case 6: return 2.1 * u0;
    case 7: return 2.3 * u0;
    default: return 0.0;
  }
}

/* Satellite 0 at table row `row` = g*OrbitRes + k (orbit_online.cuh:39-83). */
VW_HD void vw_lunar_sat0(int64_t row, int64_t orbit_res, int day, double* x, double* y,
                         double* z) {
  int64_t g = row / orbit_res;
  int64_t k = row - g * orbit_res;
  double raan_deg = ((double)g * 0.08 + ((double)(k + 1) / (double)orbit_res) * 0.08) +
                    (double)(day - 1) * 0.8;
  double s0 = (1.0 / (double)orbit_res) * VW_INV_TWOPI;
  double m_turn = s0;
  if (orbit_res > 1)
    m_turn = s0 + (1.0025 - s0) * ((double)k / (double)(orbit_res - 1));
  double sw, cw, sr, cr, si, ci;
  vw_sincos_turn(m_turn, &sw, &cw);
  vw_sincos_turn(raan_deg / 360.0, &sr, &cr);
  vw_sincos_turn(30.0 / 360.0, &si, &ci);
  vw_orbit_xyz(VW_LUNAR_A_KM, sw, cw, sr, cr, si, ci, x, y, z);
}

/* Along-track spacing triangle (orbit_online.cuh:111-123). */
VW_HD double vw_lunar_triangle(double r2, double r1, int64_t p, int64_t T) {
  if (T <= 1) return r2;
  if (p < T) {
    double t = (double)p / (double)(T - 1);
    return r2 + (r1 - r2) * t;
  }
  double t = (double)(p - T) / (double)(T - 1);
  return r1 + (r2 - r1) * t;
}

/* Satellite `sat` at table row `row` (orbit_online.cuh:125-161).  Row 0 of
 * satellites 1..7 is never written by the reference and stays at the origin. */
VW_HD void vw_lunar_sat(int64_t row, int sat, int64_t orbit_res, int day, int64_t tw,
                        int64_t day_offset, double* x, double* y, double* z) {
  double x0, y0, z0;
  vw_lunar_sat0(row, orbit_res, day, &x0, &y0, &z0);
  if (sat == 0) { *x = x0; *y = y0; *z = z0; return; }
  if (row == 0) { *x = 0.0; *y = 0.0; *z = 0.0; return; }
  double xp, yp, zp;
  vw_lunar_sat0(row - 1, orbit_res, day, &xp, &yp, &zp);
  double dx = x0 - xp, dy = y0 - yp, dz = z0 - zp;
  double mv = sqrt(vw_dot3(dx, dy, dz, dx, dy, dz)) + 1e-20;
  int64_t period = 2 * tw;
  int64_t p = period > 0 ? (day_offset + (row - 1)) % period : 0;
  double rt = vw_lunar_triangle(vw_lunar_r2(sat), vw_lunar_r1(sat), p, tw);
  *x = x0 - ((rt * dx) / mv) / 1000.0;
  *y = y0 - ((rt * dy) / mv) / 1000.0;
  *z = z0 - ((rt * dz) / mv) / 1000.0;
}

/* Circular orbit at time t for precomputed per-antenna angles (all turns):
 * theta = phase + rate * t. */
VW_HD void vw_circ_orbit(double radius, double phase_turn, double rate_turn, double t,
                         double sr, double cr, double si, double ci, double* x, double* y,
                         double* z) {
  double th = phase_turn + rate_turn * t;
  double sw, cw;
  vw_sincos_turn(th, &sw, &cw);
  vw_orbit_xyz(radius, sw, cw, sr, cr, si, ci, x, y, z);
}

/* Half row of pair (m < n) at one epoch (orbit_online.cuh:186-231): the
 * reference's first (sign -1) half of each epoch block:
 *   u = -(x_n - x_m) * 1e3 / lambda, endpoint a = antenna n, endpoint b = m.
 * The second half is the exact negation with endpoints swapped. */
VW_HD double vw_uvw_comp(double pn, double pm, double lambda_m) {
  return (-(pn - pm) * 1e3) / lambda_m;
}

/* ------------------------------------------------------------------ */
/* segment parameters                                                   */
/* ------------------------------------------------------------------ */

typedef struct {
  double fa, fb, denom;
  double du, inv_du;
  int64_t res, half;
} vw_segparams;

/* Least-squares plane w ~ fa u + fb v through the origin from the five sums
 * (viss_recon_kernel.cuh:673-701) and the uv-grid resolution
 * (main_recon_online.cu:377-383). */
VW_HD void vw_segparams_from_sums(double su2, double sv2, double suv, double suw, double svw,
                                  double bl_max_m, double lambda_m, vw_segparams* sp) {
  double p = su2 * sv2;
  double denom = p - suv * suv;
  sp->denom = denom;
  double scale = (p > 0.0) ? p : 1.0;
  double eps = 1e-24 * scale;
  if (!isfinite(denom) || fabs(denom) < eps) {
    double sign = (denom >= 0.0) ? 1.0 : -1.0;
    denom = sign * eps;
  }
  double a = (sv2 * suw - suv * svw) / denom;
  double b = (su2 * svw - suv * suw) / denom;
  if (!isfinite(a) || !isfinite(b)) { a = 0.0; b = 0.0; }
  sp->fa = a;
  sp->fb = b;
  double lmax = sqrt(1.0 + a * a);
  double mmax = sqrt(1.0 + b * b);
  double lmmax = fmax(lmax, mmax);
  double span = (2.0 * bl_max_m) / lambda_m;
  double temp = ceil((span * 2.0) * lmmax);
  double parity = temp - 2.0 * floor(temp * 0.5);
  int64_t res = (int64_t)(((temp + 2.0) + 1.0) - parity);
  sp->res = res;
  sp->half = (res - 1) / 2;
  sp->du = span / (double)(res - 1);
  sp->inv_du = 1.0 / sp->du;
}

/* uv-grid key; 0 marks a row outside the grid.  mirror=1: floor(x+0.5)
 * (recon_mirror .cuh:462-481); mirror=0: round half away from zero
 * (viss_recon_kernel.cuh:295-312). */
VW_HD int64_t vw_grid_key(double u, double v, const vw_segparams* sp, int mirror) {
  double xu = u * sp->inv_du;
  double xv = v * sp->inv_du;
  double ru, rv;
  if (mirror) {
    ru = floor(xu + 0.5);
    rv = floor(xv + 0.5);
  } else {
    ru = (xu >= 0.0) ? floor(xu + 0.5) : ceil(xu - 0.5);
    rv = (xv >= 0.0) ? floor(xv + 0.5) : ceil(xv - 0.5);
  }
  double fu = ru + (double)sp->half;
  double fv = rv + (double)sp->half;
  double fr = (double)sp->res;
  if (!(fu >= 0.0 && fu < fr && fv >= 0.0 && fv < fr)) return 0;
  return (int64_t)fu * sp->res + (int64_t)fv + 1;
}

/* Grid cell centre in wavelengths for a non-zero key. */
VW_HD void vw_key_uv(int64_t key, const vw_segparams* sp, double* ugu, double* vgu) {
  int64_t t = key - 1;
  int64_t U = t / sp->res;
  int64_t V = t - U * sp->res;
  *ugu = (double)(U - sp->half) * sp->du;
  *vgu = (double)(V - sp->half) * sp->du;
}

/* ------------------------------------------------------------------ */
/* per-term formulas                                                    */
/* ------------------------------------------------------------------ */

/* Viss phase in turns: u l + v m + w (n-1)   (viss_recon_kernel.cuh:275) */
VW_HD double vw_viss_phase(double u, double v, double w, double l, double m, double nm1) {
  return fma(w, nm1, fma(v, m, u * l));
}

/* Vh * exp(-j 2 pi w)   (common.hpp:255-264) */
VW_HD void vw_phase_correct(double re, double im, double w, double* ore, double* oim) {
  double sw, cw;
  vw_sincos_turn(w, &sw, &cw);
  *ore = fma(re, cw, im * sw);
  *oim = fma(im, cw, -(re * sw));
}

/* Recon phase in turns: ugu (l + fa n) + vgu (m + fb n)
 * (viss_recon_kernel.cuh:554-556, 628). */
VW_HD double vw_recon_phase(double ugu, double vgu, double lp, double mp) {
  return fma(vgu, mp, ugu * lp);
}

/* acc + Re(V exp(+j 2 pi x)) = acc + Vr cos - Vi sin, the phase passed as
 * t = 1024 x; the sine product is added first */
VW_HD double vw_recon_add(double acc, double vr, double vi, double s, double c) {
  return fma(vr, c, fma(-vi, s, acc));
}

/* lm-grid image point (row i -> m, column j -> l), pixel centres on [-1,1]:
 * l = (2j+1-N)/N, m = (2i+1-N)/N, n = +sqrt(1-l^2-m^2); inside the unit disk
 * iff (2j+1-N)^2 + (2i+1-N)^2 < N^2 (exact integer test). */
VW_HD int vw_grid_point(int64_t N, int64_t i, int64_t j, double* l, double* m, double* n) {
  int64_t a = 2 * j + 1 - N;
  int64_t b = 2 * i + 1 - N;
  int64_t r2 = a * a + b * b;
  *l = (double)a / (double)N;
  *m = (double)b / (double)N;
  if (r2 >= N * N) { *n = 0.0; return 0; }
  *n = sqrt((double)(N * N - r2)) / (double)N;
  return 1;
}

#endif /* VW_SHARED_H */
