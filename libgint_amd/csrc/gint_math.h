/*
 * Scalar kernels of the K-build that both the host engine and the gfx950
 * kernels evaluate: exponential, minimum-image wrap, Boys function and the
 * truncated-Coulomb G_m(R,T).
 *
 * The value of Fm[0] for a (primitive quartet, cell) slot decides whether the
 * slot runs the recurrences (Fm[0] > 1e-12).  To make that decision the same
 * on every device, every function here uses only correctly rounded IEEE
 * operations (+ - * /, sqrt, rint, ldexp), no fused multiply-add, and a fixed
 * evaluation order; FP contraction is off inside every function here.
 */
#ifndef LIBGINT_GINT_MATH_H
#define LIBGINT_GINT_MATH_H

#include <math.h>
#include "tcg_patches.h"

#if defined(__HIPCC__)
#define GINT_HD __host__ __device__ inline
#else
#define GINT_HD static inline
#endif

/* first statement of every function below: no FMA contraction in its body
   (host code built by gcc uses -ffp-contract=off for the same effect) */
#if defined(__clang__)
#define GINT_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define GINT_NO_CONTRACT
#endif

/* 2 pi^(5/2), the Obara-Saika (ss|ss) prefactor numerator */
#define GINT_TWO_PI_52 34.986836655249725
#define GINT_SQRT2     1.4142135623730950488
#define GINT_PI        3.14159265358979323846

/*
 * exp(x): Cody-Waite reduction x = k ln2 + r with ln2 split so that k*LN2_HI
 * is exact (|k| < 2^11), |r| <= ln2/2, then the degree-13 Taylor polynomial
 * of e^r in Horner form and an exact ldexp.  Error below 2 ulp.
 */
GINT_HD double gint_exp(double x)
{
   GINT_NO_CONTRACT
   const double LOG2E = 1.4426950408889634074;
   const double LN2_HI = 6.93147180369123816490e-01;
   const double LN2_LO = 1.90821492927058770002e-10;
   if (!(x > -745.2)) return (x != x) ? x : 0.0;
   if (x > 709.78) return HUGE_VAL;
   double kd = rint(x * LOG2E);
   double r = (x - kd * LN2_HI) - kd * LN2_LO;
   double p = 1.6059043836821614599e-10;          /* 1/13! */
   p = p * r + 2.0876756987868098979e-09;         /* 1/12! */
   p = p * r + 2.5052108385441718775e-08;         /* 1/11! */
   p = p * r + 2.7557319223985890653e-07;         /* 1/10! */
   p = p * r + 2.7557319223985890653e-06;         /* 1/9!  */
   p = p * r + 2.4801587301587301587e-05;         /* 1/8!  */
   p = p * r + 1.9841269841269841270e-04;         /* 1/7!  */
   p = p * r + 1.3888888888888888889e-03;         /* 1/6!  */
   p = p * r + 8.3333333333333333333e-03;         /* 1/5!  */
   p = p * r + 4.1666666666666666667e-02;         /* 1/4!  */
   p = p * r + 1.6666666666666666667e-01;         /* 1/3!  */
   p = p * r + 0.5;
   p = p * r + 1.0;
   p = p * r + 1.0;
   return ldexp(p, (int)kd);
}

/*
 * AB = B - A wrapped to the minimum image of the cell (reference
 * src/util.cpp:132-162).  cell[0..8] is the h matrix, cell[9..17] its
 * inverse, both column-major with lattice vectors as columns.
 */
GINT_HD void gint_pbc(const double A[3], const double B[3], const double *cell, double AB[3])
{
   GINT_NO_CONTRACT
   const double *h = cell;
   const double *hi = cell + 9;
   double d0 = B[0] - A[0], d1 = B[1] - A[1], d2 = B[2] - A[2];
   double s0 = hi[0] * d0 + hi[3] * d1 + hi[6] * d2;
   double s1 = hi[1] * d0 + hi[4] * d1 + hi[7] * d2;
   double s2 = hi[2] * d0 + hi[5] * d1 + hi[8] * d2;
   s0 -= round(s0);
   s1 -= round(s1);
   s2 -= round(s2);
   AB[0] = h[0] * s0 + h[3] * s1 + h[6] * s2;
   AB[1] = h[1] * s0 + h[4] * s1 + h[7] * s2;
   AB[2] = h[2] * s0 + h[5] * s1 + h[8] * s2;
}

/*
 * Boys function F_n(T), n = 0..nmax (nmax <= 21), times fac.  Reference
 * src/fgamma.cpp:196-243: T < 1e-13 exact limit; T < 12 six-term Taylor
 * expansion about the nearest 0.1 grid point of the table (121 x 28,
 * produced by gint_boys_table) followed by downward recursion; otherwise the
 * large-T formula with a four-term correction and upward recursion.
 */
GINT_HD void gint_boys(int nmax, double T, double *f, const double *ftable, double fac)
{
   GINT_NO_CONTRACT
   if (T < 1.e-13) {
      for (int n = 0; n <= nmax; n++) f[n] = fac / (2. * n + 1.);
      return;
   }
   if (T < 12.) {
      const double tdelta = 0.1;
      const int ld = 28;
      int itab = (int)round(T / tdelta);
      double ttab = (double)itab * tdelta;
      const double *row = ftable + itab * ld + nmax;
      double dt = ttab - T;
      double tmp = dt;
      double fL = row[0];
      fL += row[1] * tmp * 1.0;
      tmp *= dt; fL += row[2] * tmp * 0.5;
      tmp *= dt; fL += row[3] * tmp * (1. / 6.);
      tmp *= dt; fL += row[4] * tmp * (1. / 24.);
      tmp *= dt; fL += row[5] * tmp * (1. / 120.);
      tmp *= dt; fL += row[6] * tmp * (1. / 720.);
      double expt = gint_exp(-T);
      f[nmax] = fL;
      for (int n = nmax - 1; n >= 0; n--) f[n] = (2.0 * T * f[n + 1] + expt) / (2. * n + 1.);
   } else {
      double tmp = 1.0 / T;
      double tmp2 = tmp * tmp;
      double tmp3 = tmp * tmp2;
      double g = 0.4999489092 - 0.2473631686 * tmp + 0.321180909 * tmp2 - 0.3811559346 * tmp3;
      double expt = gint_exp(-T);
      f[0] = 0.5 * sqrt(GINT_PI * tmp) - g * expt * tmp;
      for (int n = 1; n <= nmax; n++) f[n] = (0.5 * tmp) * ((2. * n - 1.) * f[n - 1] - expt);
   }
   for (int n = 0; n <= nmax; n++) f[n] *= fac;
}

/* Where G_m(R,T) comes from for one slot. */
#define GINT_G_ZERO  0   /* beyond truncation: all G_m = 0            */
#define GINT_G_BOYS  1   /* cutoff irrelevant: plain Boys function     */
#define GINT_G_PATCH 2   /* interpolate patch `patch` at (TG1,TG2)     */

/*
 * Classify (R,T) and map it into its patch (reference
 * src/t_c_g0_n.cpp:78-145).  grid_lo / grid_hi are the 128 x 128 patch
 * indices of the low-R / high-R squares and bw the per-patch bias/width
 * table B1 W1 B2 W2 (both from gint_fill_patch_tables).
 */
GINT_HD int gint_tcg_locate(double R, double T, const int *grid_lo, const int *grid_hi,
                            const double *bw, int *patch, double *TG1, double *TG2)
{
   GINT_NO_CONTRACT
   double upper = R * R + 11.0 * R + 50.0;
   double lower = R * R - 11.0 * R + 0.0;
   double X1, X2;
   int p;
   if (T > upper) return GINT_G_ZERO;
   if (R <= 11.0) {
      X2 = R / 11.0;
      upper = R * R + 11.0 * R + 50.0;
      lower = 0.0;
      X1 = (T - lower) / (upper - lower);
      int i1 = (int)(X1 * TCG_GRID);
      int i2 = (int)(X2 * TCG_GRID);
      if (i1 == TCG_GRID) i1 = TCG_GRID - 1;
      if (i2 == TCG_GRID) i2 = TCG_GRID - 1;
      p = grid_lo[i1 * TCG_GRID + i2];
   } else {
      if (T < lower) return GINT_G_BOYS;
      X2 = 11.0 / R;
      X1 = (T - lower) / (upper - lower);
      int i1 = (int)(X1 * TCG_GRID);
      int i2 = (int)(X2 * TCG_GRID);
      if (i1 == TCG_GRID) i1 = TCG_GRID - 1;
      if (i2 == TCG_GRID) i2 = TCG_GRID - 1;
      p = grid_hi[i1 * TCG_GRID + i2];
   }
   const double B1 = bw[4 * p + 0], W1 = bw[4 * p + 1];
   const double B2 = bw[4 * p + 2], W2 = bw[4 * p + 3];
   *TG1 = (2. * X1 - B1) * W1;
   *TG2 = (2. * X2 - B2) * W2;
   *patch = p;
   return GINT_G_PATCH;
}

/* scaled Chebyshev values T_0 = 1, T_n = sqrt(2) cos(n acos x), n < 14 */
GINT_HD void gint_cheb14(double x, double *t)
{
   GINT_NO_CONTRACT
   t[0] = 1.0;
   t[1] = GINT_SQRT2 * x;
   t[2] = 2. * x * t[1] - GINT_SQRT2;
   for (int i = 3; i < 14; i++) t[i] = 2. * x * t[i - 1] - t[i - 2];
}

/* one derivative order of a patch expansion, in the fixed order of
   reference src/t_c_g0_n.cpp:163-178 */
GINT_HD double gint_cheb_sum(const double *c105, const double *T1, const double *T2)
{
   GINT_NO_CONTRACT
   double tmp = 0.0;
   int jl = 0;
   for (int j = 0; j < 14; j++) {
      double dot = 0.0;
      for (int l = 0; l < 14 - j; l++) {
         dot += T1[l] * c105[jl];
         jl++;
      }
      tmp += dot * T2[j];
   }
   return tmp;
}

/*
 * Host-side set-up of the patch lookup: the 128 x 128 grids (cell (i1,i2)
 * belongs to the box that contains it; boxes are unions of whole cells;
 * reference src/t_c_g0_n.cpp:53-64 builds the same grids from its decision
 * trees) and the bias/width table TG = (2X - B) W of each box (reference
 * src/BW_by_patch.h).  All values are exact dyadic numbers.
 */
static inline void gint_fill_patch_tables(int *grid_lo, int *grid_hi, double *bw)
{
   for (int p = 0; p < TCG_NPATCH; p++) {
      const unsigned char *bx = tcg_patch_box[p];
      int *g = (p < TCG_NPATCH_LOW) ? grid_lo : grid_hi;
      for (int i1 = bx[0]; i1 < bx[1]; i1++)
         for (int i2 = bx[2]; i2 < bx[3]; i2++) g[i1 * TCG_GRID + i2] = p;
      bw[4 * p + 0] = (double)(bx[0] + bx[1]) / 128.0;
      bw[4 * p + 1] = 128.0 / (double)(bx[1] - bx[0]);
      bw[4 * p + 2] = (double)(bx[2] + bx[3]) / 128.0;
      bw[4 * p + 3] = 128.0 / (double)(bx[3] - bx[2]);
   }
}

#endif
