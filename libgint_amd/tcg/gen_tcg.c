/*
 * gen_tcg.c -- builds the coefficient table ("C0") of the truncated-Coulomb
 * auxiliary function used by the K-build, in the layout CP2K's t_c_g.dat
 * files use, because that data file is not shipped with the library.
 *
 * Definition.  For a Gaussian charge pair with reduced exponent rho and
 * separation |PQ|, T = rho |PQ|^2 and R = R_cut sqrt(rho).  The truncated
 * Coulomb kernel 1/r (r < R_cut, 0 beyond) gives
 *
 *   G_0(R,T) = sqrt(pi)/(4 sqrt T) [ 2 erf(sqrt T) + erf(R - sqrt T) - erf(R + sqrt T) ]
 *   G_m(R,T) = (-d/dT)^m G_0(R,T)          (R held fixed),
 *
 * with G_m -> F_m (the Boys function) as R -> infinity.  Reference
 * src/t_c_g0_n.cpp:67-180 consumes exactly this function, scaled by the
 * usual Obara-Saika prefactor.
 *
 * Evaluation (all in IEEE binary128 via libquadmath):
 *   T <= T_SWITCH : Taylor series about T = 0,
 *                   G_m(R,T) = sum_n G_{m+n}(R,0) (-T)^n / n!,
 *                   G_j(R,0) = 1/(2j+1) - (-1)^j H_{2j}(R) e^{-R^2} / (2^j (2j+1)!!)
 *                   (H = physicists' Hermite polynomial; derived from the
 *                   Fourier form G_m = pi^-1/2 Int_0^inf (1-cos kR) e^{-k^2/4}
 *                   (k^2/2)^m j_m(k sqrt T)/(k sqrt T)^m dk).
 *   T >  T_SWITCH : truncated Taylor-jet arithmetic on the closed form of G_0
 *                   about T, G_m = (-1)^m m! [jet]_m.
 *
 * Fit.  Patch p covers X1 in [a1,b1]/128, X2 in [a2,b2]/128 (tcg_patches.h).
 * (X1,X2) map to (R,T) exactly as the lookup does (reference
 * src/t_c_g0_n.cpp:90-131):
 *   low R  : R = 11 X2,   T = X1 (R^2 + 11 R + 50)
 *   high R : R = 11 / X2, T = (R^2 - 11 R) + X1 (22 R + 50)
 * and TG = (2X - B) W maps each box to [-1,1]^2.  The expansion uses the
 * scaled Chebyshev basis T_0 = 1, T_n = sqrt(2) cos(n theta) (orthonormal
 * under the discrete Gauss-Chebyshev rule), truncated to l + j <= 13, and
 * is stored as C0[p][k*105 + jl] with jl running j = 0..13 (outer),
 * l = 0..13-j (inner), the order src/t_c_g0_n.cpp:163-178 reads it in.
 * Coefficients are the discrete Chebyshev transform on an N x N grid.
 *
 * Usage: gen_tcg OUT.bin [NDERIV]   (writes 207*(NDERIV+1)*105 doubles,
 * little-endian, patch-major; prints a fit-error summary on stdout.)
 */
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "../csrc/tcg_patches.h"

typedef __float128 fq;

#define NJ        (TCG_NDERIV_MAX + 1)   /* derivative orders kept in a jet */
#define NSERIES   48                     /* terms of the T = 0 series     */
#define T_SWITCH  3.0Q
#define NGRID     32                     /* Gauss-Chebyshev nodes per axis */
#define NTEST     256                    /* random check points per patch  */

static const fq QPI = 3.141592653589793238462643383279502884Q;

/* G_j(R,0), j = 0..J */
static void g_at_zero(fq R, int J, fq *g)
{
   fq h[2 * (TCG_NDERIV_MAX + NSERIES) + 4];
   const int nmax = 2 * J;
   h[0] = expq(-R * R);
   h[1] = 2 * R * h[0];
   for (int n = 1; n < nmax; n++) h[n + 1] = 2 * R * h[n] - 2 * n * h[n - 1];
   fq den = 1;
   for (int j = 0; j <= J; j++) {
      if (j > 0) den *= 2 * (2 * j + 1);
      fq sgn = (j & 1) ? -1 : 1;
      g[j] = 1 / (fq)(2 * j + 1) - sgn * h[2 * j] / den;
   }
}

/* ---- truncated Taylor jets in d = T - T0, NJ coefficients ---- */
static void jet_mul(const fq *a, const fq *b, fq *c)
{
   for (int k = 0; k < NJ; k++) {
      fq s = 0;
      for (int j = 0; j <= k; j++) s += a[j] * b[k - j];
      c[k] = s;
   }
}
static void jet_div(const fq *a, const fq *b, fq *c)
{
   for (int k = 0; k < NJ; k++) {
      fq s = a[k];
      for (int j = 1; j <= k; j++) s -= b[j] * c[k - j];
      c[k] = s / b[0];
   }
}
static void jet_exp(const fq *u, fq *y)
{
   y[0] = expq(u[0]);
   for (int k = 1; k < NJ; k++) {
      fq s = 0;
      for (int j = 1; j <= k; j++) s += j * u[j] * y[k - j];
      y[k] = s / k;
   }
}
/* erf of a jet x */
static void jet_erf(const fq *x, fq *e)
{
   fq x2[NJ], mx2[NJ], w[NJ], xp[NJ], p[NJ];
   jet_mul(x, x, x2);
   for (int k = 0; k < NJ; k++) mx2[k] = -x2[k];
   jet_exp(mx2, w);
   for (int k = 0; k < NJ; k++) xp[k] = (k + 1 < NJ) ? (k + 1) * x[k + 1] : 0;
   jet_mul(w, xp, p);
   const fq c = 2 / sqrtq(QPI);
   e[0] = erfq(x[0]);
   for (int k = 1; k < NJ; k++) e[k] = c * p[k - 1] / k;
}

/* G_m(R,T), m = 0..NJ-1 */
static void g_eval(fq R, fq T, fq *G)
{
   if (T <= T_SWITCH) {
      fq g0[TCG_NDERIV_MAX + NSERIES + 1];
      g_at_zero(R, TCG_NDERIV_MAX + NSERIES, g0);
      for (int m = 0; m < NJ; m++) {
         fq s = 0, t = 1;
         for (int n = 0; n <= NSERIES; n++) {
            s += g0[m + n] * t;
            t *= -T / (n + 1);
         }
         G[m] = s;
      }
      return;
   }
   fq sj[NJ], a[NJ], b[NJ], ea[NJ], eb[NJ], es[NJ], br[NJ], res[NJ];
   /* sqrt(T0 + d) = sqrt(T0) sum_k binom(1/2,k) (d/T0)^k */
   fq s0 = sqrtq(T), bin = 1;
   for (int k = 0; k < NJ; k++) {
      sj[k] = s0 * bin;
      bin *= (0.5Q - k) / (k + 1) / T;
   }
   for (int k = 0; k < NJ; k++) { a[k] = -sj[k]; b[k] = sj[k]; }
   a[0] += R; b[0] += R;
   jet_erf(sj, es);
   jet_erf(a, ea);
   jet_erf(b, eb);
   for (int k = 0; k < NJ; k++) br[k] = sqrtq(QPI) / 4 * (2 * es[k] + ea[k] - eb[k]);
   jet_div(br, sj, res);
   fq f = 1;
   for (int m = 0; m < NJ; m++) {
      if (m > 0) f *= m;
      G[m] = ((m & 1) ? -f : f) * res[m];
   }
}

static void patch_RT(int p, fq X1, fq X2, fq *R, fq *T)
{
   if (p < TCG_NPATCH_LOW) {
      *R = 11 * X2;
      *T = X1 * ((*R) * (*R) + 11 * (*R) + 50);
   } else {
      *R = 11 / X2;
      fq lo = (*R) * (*R) - 11 * (*R);
      *T = lo + X1 * (22 * (*R) + 50);
   }
}

/* evaluation of one patch expansion in double, in the order of
   reference src/t_c_g0_n.cpp:144-178 */
static double eval_double(const double *row, int k, double X1, double X2, int p)
{
   const unsigned char *bx = tcg_patch_box[p];
   double B1 = (bx[0] + bx[1]) / 128.0, W1 = 128.0 / (bx[1] - bx[0]);
   double B2 = (bx[2] + bx[3]) / 128.0, W2 = 128.0 / (bx[3] - bx[2]);
   double TG1 = (2. * X1 - B1) * W1, TG2 = (2. * X2 - B2) * W2;
   const double SQ2 = 1.4142135623730950488;
   double T1[14], T2[14];
   T1[0] = 1.0; T2[0] = 1.0;
   T1[1] = SQ2 * TG1; T2[1] = SQ2 * TG2;
   T1[2] = 2. * TG1 * T1[1] - SQ2; T2[2] = 2. * TG2 * T2[1] - SQ2;
   for (int i = 3; i < 14; i++) {
      T1[i] = 2. * TG1 * T1[i - 1] - T1[i - 2];
      T2[i] = 2. * TG2 * T2[i - 1] - T2[i - 2];
   }
   double tmp = 0.0;
   int jl = 0;
   for (int j = 0; j < 14; j++) {
      double dot = 0.0;
      for (int l = 0; l < 14 - j; l++) { dot += T1[l] * row[k * 105 + jl]; jl++; }
      tmp += dot * T2[j];
   }
   return tmp;
}

static unsigned long long lcg(unsigned long long *s)
{
   *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
   return *s >> 11;
}

int main(int argc, char **argv)
{
   if (argc < 2) { fprintf(stderr, "usage: %s OUT.bin [NDERIV] | --eval R T\n", argv[0]); return 2; }
   if (strcmp(argv[1], "--eval") == 0 && argc == 4) {
      fq G[NJ];
      g_eval(strtoflt128(argv[2], NULL), strtoflt128(argv[3], NULL), G);
      for (int m = 0; m < NJ; m++) {
         char buf[64];
         quadmath_snprintf(buf, sizeof buf, "%.30Qe", G[m]);
         printf("%s\n", buf);
      }
      return 0;
   }
   int nder = (argc > 2) ? atoi(argv[2]) : TCG_NDERIV_MAX;
   if (nder < 0 || nder > TCG_NDERIV_MAX) { fprintf(stderr, "NDERIV out of range\n"); return 2; }
   const int ld = (nder + 1) * TCG_NCOEF;
   double *C0 = calloc((size_t)TCG_NPATCH * ld, sizeof(double));
   double err_abs[NJ] = {0}, err_rel[NJ] = {0};

   #pragma omp parallel for schedule(dynamic, 1)
   for (int p = 0; p < TCG_NPATCH; p++) {
      const unsigned char *bx = tcg_patch_box[p];
      static __thread fq F[NGRID][NGRID][NJ];
      fq xn[NGRID], th[NGRID];
      for (int i = 0; i < NGRID; i++) {
         th[i] = QPI * (i + 0.5Q) / NGRID;
         xn[i] = cosq(th[i]);
      }
      for (int i1 = 0; i1 < NGRID; i1++)
         for (int i2 = 0; i2 < NGRID; i2++) {
            fq X1 = (bx[0] + bx[1] + xn[i1] * (bx[1] - bx[0])) / 256;
            fq X2 = (bx[2] + bx[3] + xn[i2] * (bx[3] - bx[2])) / 256;
            fq R, T;
            patch_RT(p, X1, X2, &R, &T);
            g_eval(R, T, F[i1][i2]);
         }
      /* basis values */
      fq Tb[NGRID][TCG_DEGREE + 1];
      for (int i = 0; i < NGRID; i++)
         for (int n = 0; n <= TCG_DEGREE; n++)
            Tb[i][n] = (n == 0) ? 1 : sqrtq(2.0Q) * cosq(n * th[i]);
      for (int k = 0; k <= nder; k++) {
         int jl = 0;
         for (int j = 0; j <= TCG_DEGREE; j++)
            for (int l = 0; l <= TCG_DEGREE - j; l++) {
               fq s = 0;
               for (int i1 = 0; i1 < NGRID; i1++) {
                  fq s2 = 0;
                  for (int i2 = 0; i2 < NGRID; i2++) s2 += F[i1][i2][k] * Tb[i2][j];
                  s += s2 * Tb[i1][l];
               }
               C0[(size_t)p * ld + k * TCG_NCOEF + jl] = (double)(s / (NGRID * NGRID));
               jl++;
            }
      }
      /* fit error at random interior points */
      unsigned long long seed = 12345 + 7919ULL * p;
      double ea[NJ] = {0}, er[NJ] = {0};
      for (int t = 0; t < NTEST; t++) {
         double u1 = (lcg(&seed) + 0.5) / 9007199254740992.0;
         double u2 = (lcg(&seed) + 0.5) / 9007199254740992.0;
         double X1 = (bx[0] + u1 * (bx[1] - bx[0])) / 128.0;
         double X2 = (bx[2] + u2 * (bx[3] - bx[2])) / 128.0;
         fq R, T, G[NJ];
         patch_RT(p, X1, X2, &R, &T);
         g_eval(R, T, G);
         for (int k = 0; k <= nder; k++) {
            double v = eval_double(&C0[(size_t)p * ld], k, X1, X2, p);
            double d = fabs(v - (double)G[k]);
            double scale = fabs((double)G[k]);
            if (d > ea[k]) ea[k] = d;
            if (scale > 1e-300 && d / scale > er[k]) er[k] = d / scale;
         }
      }
      #pragma omp critical
      for (int k = 0; k <= nder; k++) {
         if (ea[k] > err_abs[k]) err_abs[k] = ea[k];
         if (er[k] > err_rel[k]) err_rel[k] = er[k];
      }
   }
   FILE *f = fopen(argv[1], "wb");
   if (!f) { perror(argv[1]); return 1; }
   size_t n = fwrite(C0, sizeof(double), (size_t)TCG_NPATCH * ld, f);
   fclose(f);
   if (n != (size_t)TCG_NPATCH * ld) { fprintf(stderr, "short write\n"); return 1; }
   printf("nderiv %d ld_C0 %d patches %d\n", nder, ld, TCG_NPATCH);
   for (int k = 0; k <= nder; k++)
      printf("m %2d  max_abs_err %.3e  max_rel_err %.3e\n", k, err_abs[k], err_rel[k]);
   free(C0);
   return 0;
}
