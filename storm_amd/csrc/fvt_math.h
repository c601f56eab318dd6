// Numerical fluxes and small helpers of the FVT step, shared by the device
// kernels.  Expression order follows the reference routines cited per
// function so that results agree with the reference arithmetic.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

namespace storm {

#define SD __device__ __forceinline__

constexpr double G = 9.81;
constexpr double HALF = 0.5;
constexpr double ONE = 1.0;
constexpr double ZERO = 0.0;
constexpr double ONE_THIRD = 1.0 / 3.0;
// REAL(mp), PARAMETER :: small = 1.0E-9 is a default-real literal
// (flux_ac.f90:46, solverFVT.f90:44, deltatFVT.f90:30)
constexpr double SMALL = (double)1.0e-9f;

// h**one_third of the Manning friction factor (fric_terms_impl.f90:36,
// fric_terms.f90:37), correctly rounded.  The exponent is fl(1/3) =
// 1/3 - 2^-54/3, so h**fl(1/3) = cbrt(h) * (1 - (2^-54/3) ln h) to far below
// an ulp.  The cube root starts from single precision: h = m 2^(3q) with m in
// [0.5, 4), cbrt(m) = 2^(log2(m) / 3) from the hardware's single-precision
// log2 and exp2 (relative error near 1e-7), one plain Newton step in double
// (error near 1e-14), then one Newton step whose residual c^3 - h is formed
// with exact products (fma), which leaves the root good to about 2^-90; the
// exponent's correction (its ln h from the same log2, good to 1e-7, which is
// ample for a term of 2^-54 relative size) and that step are added in one
// final rounding.  The result is the correctly rounded value on all but
// inputs within about 2^-20 ulp of a rounding midpoint.  A libm pow is not
// correctly rounded: glibc's (0.52 ulp bound) differs on about 8 inputs in
// 10,000, the device's own more often, and the step amplifies such one-ulp
// differences over thousands of steps.  h <= 0 or not finite: plain pow.
SD double pow_one_third(double h) {
  if (!(h > 0.0 && h <= DBL_MAX)) return pow(h, 1.0 / 3.0);
  int e;
  const double m = frexp(h, &e);               // h = m 2^e, m in [0.5, 1)
  const int q = (e >= 0 ? e : e - 2) / 3;      // floor(e / 3)
  const double mr = ldexp(m, e - 3 * q);       // h = mr 2^(3q), mr in [0.5, 4)
  const float lf = __builtin_amdgcn_logf((float)mr);  // log2(mr)
  double c = ldexp((double)__builtin_amdgcn_exp2f(lf * (1.0f / 3.0f)), q);
  c = c - (c * c * c - h) / (3.0 * (c * c));
  const double p = c * c, pe = fma(c, c, -p);   // c^2 = p + pe exactly
  const double pc = p * c, qe = fma(p, c, -pc); // p c = pc + qe exactly
  const double r = (pc - h) + (qe + pe * c);    // c^3 - h
  const double lo = -r / (3.0 * p);             // Newton correction
  const double lnh = ((double)lf + 3.0 * q) * 0.6931471805599453;
  const double t = -1.8503717077085943e-17 * lnh;  // -(1/3 - fl(1/3)) ln h
  return c + (lo + c * t);
}

// Quotients that share a denominator d, each with the bits of the plain
// IEEE division n / d.  The compiler's gfx950 division is: d scaled by
// v_div_scale (a no-op unless n / d is near overflow or underflow), its
// reciprocal refined by two Newton steps, n scaled, q = n * r, one
// correction fma, v_div_fmas and v_div_fixup.  `recip` runs the part that
// depends on d alone once; `qdiv` runs the rest per numerator, after checking
// that this numerator scales d the same way (otherwise it divides the plain
// way).  Same instructions on the same operands, so the same bits.
struct Recip {
  double ds, r;  // scaled denominator, refined reciprocal
};
SD Recip recip(double d, double n) {
  bool f;
  const double ds = __builtin_amdgcn_div_scale(n, d, false, &f);
  double r = __builtin_amdgcn_rcp(ds);
  double e = fma(-ds, r, 1.0);
  r = fma(r, e, r);
  e = fma(-ds, r, 1.0);
  r = fma(r, e, r);
  return Recip{ds, r};
}
SD double qdiv(const Recip& R, double n, double d) {
  bool f, scaled;
  const double ds = __builtin_amdgcn_div_scale(n, d, false, &f);
  if (__double_as_longlong(ds) != __double_as_longlong(R.ds)) return n / d;
  const double ns = __builtin_amdgcn_div_scale(n, d, true, &scaled);
  const double q = ns * R.r;
  const double rem = fma(-R.ds, q, ns);
  return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(rem, R.r, q, scaled), d, n);
}

SD double dmin(double a, double b) { return (a < b) ? a : b; }
SD double dmax(double a, double b) { return (a > b) ? a : b; }

// equals.f90:14-24: relative-epsilon equality
SD bool equals(double a, double b) {
  double eps = fabs(a) * DBL_EPSILON;
  if (eps == ZERO) eps = DBL_MIN;
  return !(fabs(a - b) > eps);
}

// flux_ac.f90:48-126: Roe flux of Anastasiou & Chan (1997).  The 3x3
// products R*|Lambda|*L*dq are formed as in the reference (MATMUL chain, sums
// in index order); R*|Lambda| needs no sums because |Lambda| is diagonal.
SD void flux_ac(double nx, double ny, double hl, double ul, double vl, double hr, double ur,
                double vr, double* f) {
  const double delta = 0.10;
  const double fl0 = hl * ul, fl1 = hl * ul * ul + HALF * G * hl * hl, fl2 = hl * ul * vl;
  const double gl0 = hl * vl, gl1 = hl * ul * vl, gl2 = hl * vl * vl + HALF * G * hl * hl;
  const double fr0 = hr * ur, fr1 = hr * ur * ur + HALF * G * hr * hr, fr2 = hr * ur * vr;
  const double gr0 = hr * vr, gr1 = hr * ur * vr, gr2 = hr * vr * vr + HALF * G * hr * hr;
  hl = dmax(hl, SMALL);
  const double hls = sqrt(hl);
  hr = dmax(hr, SMALL);
  const double hrs = sqrt(hr);
  const double den = hrs + hls;
  const double nua = ur * hrs + ul * hls, nva = vr * hrs + vl * hls;
  const Recip rden = recip(den, nua);
  const double ua = qdiv(rden, nua, den);
  const double va = qdiv(rden, nva, den);
  const double ca = sqrt(HALF * G * (hl + hr));
  const double q0 = hr - hl, q1 = hr * ur - hl * ul, q2 = hr * vr - hl * vl;
  // right eigenvectors (columns) times |lambda|
  const double e0 = dmax(fabs(ua * nx + va * ny), delta);
  const double e1 = dmax(fabs(ua * nx + va * ny - ca), delta);
  const double e2 = dmax(fabs(ua * nx + va * ny + ca), delta);
  const double t00 = ZERO * e0, t01 = ONE * e1, t02 = ONE * e2;
  const double t10 = ny * e0, t11 = (ua - ca * nx) * e1, t12 = (ua + ca * nx) * e2;
  const double t20 = (-nx) * e0, t21 = (va - ca * ny) * e1, t22 = (va + ca * ny) * e2;
  // left eigenvectors (rows)
  // HALF * x / ca is (HALF * x) / ca: three quotients by ca
  const double n1 = HALF * (ua * nx + va * ny), n2 = HALF * nx, n3 = HALF * ny;
  const Recip rca = recip(ca, n1);
  const double c1 = qdiv(rca, n1, ca), c2 = qdiv(rca, n2, ca), c3 = qdiv(rca, n3, ca);
  const double l00 = -(ua * ny - va * nx), l01 = ny, l02 = -nx;
  const double l10 = c1 + HALF, l11 = -c2, l12 = -c3;
  const double l20 = -c1 + HALF, l21 = c2, l22 = c3;
  const double a00 = t00 * l00 + t01 * l10 + t02 * l20, a01 = t00 * l01 + t01 * l11 + t02 * l21,
               a02 = t00 * l02 + t01 * l12 + t02 * l22;
  const double a10 = t10 * l00 + t11 * l10 + t12 * l20, a11 = t10 * l01 + t11 * l11 + t12 * l21,
               a12 = t10 * l02 + t11 * l12 + t12 * l22;
  const double a20 = t20 * l00 + t21 * l10 + t22 * l20, a21 = t20 * l01 + t21 * l11 + t22 * l21,
               a22 = t20 * l02 + t21 * l12 + t22 * l22;
  const double aq0 = a00 * q0 + a01 * q1 + a02 * q2;
  const double aq1 = a10 * q0 + a11 * q1 + a12 * q2;
  const double aq2 = a20 * q0 + a21 * q1 + a22 * q2;
  f[0] = HALF * (fr0 * nx + gr0 * ny + fl0 * nx + gl0 * ny - aq0);
  f[1] = HALF * (fr1 * nx + gr1 * ny + fl1 * nx + gl1 * ny - aq1);
  f[2] = HALF * (fr2 * nx + gr2 * ny + fl2 * nx + gl2 * ny - aq2);
}

// flux_rusanov.f90:47-80: local Lax-Friedrichs, caller-supplied celerities
SD void flux_rusanov(double nx, double ny, double hl, double ul, double vl, double cl, double hr,
                     double ur, double vr, double cr, double* f) {
  const double fl0 = hl * ul, fl1 = hl * ul * ul + HALF * G * hl * hl, fl2 = hl * ul * vl;
  const double gl0 = hl * vl, gl1 = hl * ul * vl, gl2 = hl * vl * vl + HALF * G * hl * hl;
  const double fr0 = hr * ur, fr1 = hr * ur * ur + HALF * G * hr * hr, fr2 = hr * ur * vr;
  const double gr0 = hr * vr, gr1 = hr * ur * vr, gr2 = hr * vr * vr + HALF * G * hr * hr;
  const double un_l = ul * nx + vl * ny, un_r = ur * nx + vr * ny;
  const double sp = dmax(fabs(un_l) + cl, fabs(un_r) + cr);
  const double q0 = hr - hl, q1 = hr * ur - hl * ul, q2 = hr * vr - hl * vl;
  f[0] = HALF * (fr0 * nx + gr0 * ny + fl0 * nx + gl0 * ny - sp * q0);
  f[1] = HALF * (fr1 * nx + gr1 * ny + fl1 * nx + gl1 * ny - sp * q1);
  f[2] = HALF * (fr2 * nx + gr2 * ny + fl2 * nx + gl2 * ny - sp * q2);
}

// HLL (Toro 2001; Yoon & Kang 2004).  Wave speeds with the dry-bed variants
// of flux_HLL.f90:43-69; star-region flux in its textbook form (the
// reference body at :103-109 lacks the division by sR-sL and leaves flux(3)
// unset).
SD void flux_hll(double nx, double ny, double hl, double ul, double vl, double hr, double ur,
                 double vr, double h0, double* f) {
  const double cl = sqrt(G * hl), cr = sqrt(G * hr);
  const double us = HALF * ((ul + ur) * nx + (vl + vr) * ny) + cl - cr;
  const double cs = HALF * (cl + cr) + ((ul - ur) * nx + (vl - vr) * ny) / 4.0;
  double sl, sr;
  if (hl < h0) sl = ur * nx + vr * ny - 2.0 * cr;
  else if (hr < h0) sl = ul * nx + vl * ny - cl;
  else sl = dmin(ul * nx + vl * ny - cl, us - cs);
  if (hl < h0) sr = ur * nx + vr * ny + cr;
  else if (hr < h0) sr = ul * nx + vl * ny + 2.0 * cl;
  else sr = dmax(ur * nx + vr * ny + cr, us + cs);
  const double FL0 = hl * ul * nx + hl * vl * ny;
  const double FL1 = (hl * ul * ul + HALF * G * hl * hl) * nx + hl * ul * vl * ny;
  const double FL2 = hl * ul * vl * nx + (hl * vl * vl + HALF * G * hl * hl) * ny;
  const double FR0 = hr * ur * nx + hr * vr * ny;
  const double FR1 = (hr * ur * ur + HALF * G * hr * hr) * nx + hr * ur * vr * ny;
  const double FR2 = hr * ur * vr * nx + (hr * vr * vr + HALF * G * hr * hr) * ny;
  if (sl >= ZERO) { f[0] = FL0; f[1] = FL1; f[2] = FL2; }
  else if (sr <= ZERO) { f[0] = FR0; f[1] = FR1; f[2] = FR2; }
  else {
    const double d = sr - sl;
    f[0] = (sr * FL0 - sl * FR0 + sl * sr * (hr - hl)) / d;
    f[1] = (sr * FL1 - sl * FR1 + sl * sr * (hr * ur - hl * ul)) / d;
    f[2] = (sr * FL2 - sl * FR2 + sl * sr * (hr * vr - hl * vl)) / d;
  }
}

// Roe flux of Brufau & Garcia-Navarro with pre-projected physical fluxes
// (flux_bgnvc.f90:41-103) with left eigenvectors that invert the right ones
// (the reference has L(1,1) = u.n + c and L(3,3) = vavg - ny, :79,:87).
SD void flux_bgnvc(double nx, double ny, double hl, double ul, double vl, double hr, double ur,
                   double vr, double* f) {
  const double fl0 = hl * ul * nx, fl1 = (hl * ul * ul + HALF * G * hl * hl) * nx, fl2 = hl * ul * vl * nx;
  const double gl0 = hl * vl * ny, gl1 = hl * ul * vl * ny, gl2 = (hl * vl * vl + HALF * G * hl * hl) * ny;
  const double fr0 = hr * ur * nx, fr1 = (hr * ur * ur + HALF * G * hr * hr) * nx, fr2 = hr * ur * vr * nx;
  const double gr0 = hr * vr * ny, gr1 = hr * ur * vr * ny, gr2 = (hr * vr * vr + HALF * G * hr * hr) * ny;
  const double hls = sqrt(hl), hrs = sqrt(hr);
  const double den = hrs + hls;
  const double ua = (ur * hrs + ul * hls) / den;
  const double va = (vr * hrs + vl * hls) / den;
  const double ca = sqrt(HALF * G * (hl + hr));
  const double q0 = hr - hl, q1 = hr * ur - hl * ul, q2 = hr * vr - hl * vl;
  const double un = ua * nx + va * ny;
  const double r[3][3] = {{ONE, ZERO, ONE}, {ua + ca * nx, -ca * ny, ua - ca * nx}, {va + ca * ny, ca * nx, va - ca * ny}};
  double l[3][3] = {{ca - un, nx, ny}, {2.0 * (ua * ny - va * nx), -2.0 * ny, 2.0 * nx}, {un + ca, -nx, -ny}};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) l[i][j] = HALF * l[i][j] / ca;
  const double e[3] = {fabs(un + ca), fabs(un), fabs(un - ca)};
  double aq[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double t0 = r[i][0] * e[0], t1 = r[i][1] * e[1], t2 = r[i][2] * e[2];
    double a0 = t0 * l[0][0] + t1 * l[1][0] + t2 * l[2][0];
    double a1 = t0 * l[0][1] + t1 * l[1][1] + t2 * l[2][1];
    double a2 = t0 * l[0][2] + t1 * l[1][2] + t2 * l[2][2];
    aq[i] = a0 * q0 + a1 * q1 + a2 * q2;
  }
  f[0] = HALF * (fr0 + gr0 + fl0 + gl0 - aq[0]);
  f[1] = HALF * (fr1 + gr1 + fl1 + gl1 - aq[1]);
  f[2] = HALF * (fr2 + gr2 + fl2 + gl2 - aq[2]);
}

}  // namespace storm
