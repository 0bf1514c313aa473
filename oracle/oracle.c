/*
 * CPU oracle for the libGint K-build -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain C99, FP64, built with -ffp-contract=off.  It restates the reference
 * engine's algorithm function by function (citations are to files under the
 * reference repository's src/ directory) and is used only as the checker in
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Data flow, for every set quartet at add_set():
 *   per primitive quartet and lattice cell n3: prepare (images, P, Q, Kfac)
 *   -> G_m(R,T) or F_m(T) -> Obara-Saika VRR -> sum over n3
 *   -> contraction over primitives -> HRR -> sum over the (n1,n2) rows of
 *   the quartet -> cartesian-to-spherical -> exchange update of K.
 *
 * Differences from the reference that are deliberate: nl up to 4 on every
 * centre (the reference's encodings stop at 3 for three of them), no
 * hash-only deduplication of basis data, the Coulomb potential is the
 * plain Boys function when no truncated potential was set.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <assert.h>
#include <omp.h>
#include "../libgint_amd/csrc/tcg_patches.h"
#include "../libgint_amd/csrc/c2s_table.h"

#define NCO(l) (((l) + 1) * ((l) + 2) / 2)
#define MAXL 3
#define MAXLL (4 * MAXL)
#define NCLS 256

/* ------------------------------------------------------------------ */
/* scalar functions                                                    */
/* ------------------------------------------------------------------ */

/* exp by Cody-Waite reduction + degree-13 Taylor polynomial; the product
   library uses the same published method so both sides round alike.  The
   reference calls the CUDA libm exp at src/compute_Fm.cpp:123 and
   src/fgamma.cpp:222,235. */
double oracle_exp(double x)
{
   const double LOG2E = 1.4426950408889634074;
   const double LN2_HI = 6.93147180369123816490e-01;
   const double LN2_LO = 1.90821492927058770002e-10;
   if (!(x > -745.2)) return (x != x) ? x : 0.0;
   if (x > 709.78) return HUGE_VAL;
   double kd = rint(x * LOG2E);
   double r = (x - kd * LN2_HI) - kd * LN2_LO;
   static const double c[14] = {
      1.6059043836821614599e-10, 2.0876756987868098979e-09, 2.5052108385441718775e-08,
      2.7557319223985890653e-07, 2.7557319223985890653e-06, 2.4801587301587301587e-05,
      1.9841269841269841270e-04, 1.3888888888888888889e-03, 8.3333333333333333333e-03,
      4.1666666666666666667e-02, 1.6666666666666666667e-01, 0.5, 1.0, 1.0};
   double p = c[0];
   for (int i = 1; i < 14; i++) p = p * r + c[i];
   return ldexp(p, (int)kd);
}

/* minimum-image AB = B - A, reference src/util.cpp:132-162 */
static void o_pbc(const double A[3], const double B[3], const double *cell, double AB[3])
{
   const double *h = cell, *hi = cell + 9;
   AB[0] = B[0] - A[0];
   AB[1] = B[1] - A[1];
   AB[2] = B[2] - A[2];
   double s0 = hi[0] * AB[0] + hi[3] * AB[1] + hi[6] * AB[2];
   double s1 = hi[1] * AB[0] + hi[4] * AB[1] + hi[7] * AB[2];
   double s2 = hi[2] * AB[0] + hi[5] * AB[1] + hi[8] * AB[2];
   s0 -= round(s0);
   s1 -= round(s1);
   s2 -= round(s2);
   AB[0] = h[0] * s0 + h[3] * s1 + h[6] * s2;
   AB[1] = h[1] * s0 + h[4] * s1 + h[7] * s2;
   AB[2] = h[2] * s0 + h[5] * s1 + h[8] * s2;
}

/* F_n(T) reference series, reference src/fgamma.cpp:53-107 */
static void o_fgamma_ref(int nmax, double T, double *f)
{
   const int kmax = 50;
   double r[50 + 10];
   if (T < 1.e-13) {
      for (int n = 0; n <= nmax; n++) f[n] = 1.0 / (2. * n + 1.);
   } else if (T <= 50.0) {
      r[kmax + 9] = 0.0;
      for (int j = kmax + 9; j >= 1; j--) r[j - 1] = -T / ((4. * j + 2.) - T * r[j]);
      double factor = 2.0 * sinh(0.5 * T) * exp(-0.5 * T) / T;
      for (int n = 0; n <= nmax; n++) {
         double sumtot = factor / (2. * n + 1.);
         double term = 1.0;
         for (int k = 1; k <= kmax; k++) {
            term = term * (2. * n - 2. * k + 1.) / (2. * n + 2. * k + 1);
            double p = 1.0;
            for (int j = 1; j <= k; j++) p = p * r[j - 1];
            double sumterm = factor * term * p * (2. * k + 1.) / (2. * n + 1.);
            if (fabs(sumterm) < DBL_EPSILON) break;
            assert(k < kmax);
            sumtot = sumtot + sumterm;
         }
         f[n] = sumtot;
      }
   } else {
      f[0] = 0.5 * sqrt(M_PI / T);
      double expt = exp(-T);
      for (int n = 1; n <= nmax; n++) f[n] = 0.5 * ((2. * n - 1.) * f[n - 1] - expt) / T;
   }
}

/* 121 x 28 table, reference src/fgamma.cpp:109-123 with nmax = 21, tdelta
   0.1, tmax 12 (src/libGint.cpp:488-497) */
void oracle_boys_table(double *ftable)
{
   const int n = 21 + 6;
   const double tdelta = 0.1;
   const int itabmax = (int)((12. - 0.) / tdelta);
   for (int itab = 0; itab <= itabmax; itab++) {
      double T = (double)itab * tdelta;
      o_fgamma_ref(n, T, &ftable[itab * (n + 1)]);
   }
}

/* F_n(T) * fac, reference src/fgamma.cpp:196-243 */
static void o_fgamma0(int nmax, double T, double *f, const double *ftable, double fac)
{
   if (T < 1.e-13) {
      for (int n = 0; n <= nmax; n++) f[n] = fac / (2. * n + 1.);
      return;
   }
   if (T < 12.) {
      const double tdelta = 0.1;
      const int ftable_ld = 28;
      int itab = (int)round(T / tdelta);
      double ttab = (double)itab * tdelta;
      double tmp = 1.0;
      double fL = ftable[itab * ftable_ld + nmax + 0];
      tmp *= (ttab - T); fL += ftable[itab * ftable_ld + nmax + 1] * tmp * 1.0;
      tmp *= (ttab - T); fL += ftable[itab * ftable_ld + nmax + 2] * tmp * 0.5;
      tmp *= (ttab - T); fL += ftable[itab * ftable_ld + nmax + 3] * tmp * (1. / 6.);
      tmp *= (ttab - T); fL += ftable[itab * ftable_ld + nmax + 4] * tmp * (1. / 24.);
      tmp *= (ttab - T); fL += ftable[itab * ftable_ld + nmax + 5] * tmp * (1. / 120.);
      tmp *= (ttab - T); fL += ftable[itab * ftable_ld + nmax + 6] * tmp * (1. / 720.);
      double expt = oracle_exp(-T);
      f[nmax] = fL;
      for (int n = nmax - 1; n >= 0; n--) f[n] = (2.0 * T * f[n + 1] + expt) / (2. * n + 1.);
   } else {
      double tmp = 1.0 / T;
      double tmp2 = tmp * tmp;
      double tmp3 = tmp * tmp2;
      double g = 0.4999489092 - 0.2473631686 * tmp + 0.321180909 * tmp2 - 0.3811559346 * tmp3;
      double expt = oracle_exp(-T);
      f[0] = 0.5 * sqrt(M_PI * tmp) - g * expt * tmp;
      for (int n = 1; n <= nmax; n++) f[n] = (0.5 * tmp) * ((2. * n - 1.) * f[n - 1] - expt);
   }
   for (int n = 0; n <= nmax; n++) f[n] *= fac;
}

static double *s_ftable(void)
{
   static double ft[121 * 28];
   static int done = 0;
   if (!done) { oracle_boys_table(ft); done = 1; }
   return ft;
}

void oracle_boys(int nmax, double T, double *f, double fac) { o_fgamma0(nmax, T, f, s_ftable(), fac); }

/* patch grids from the box table: grid cell i lies in box [a,b) along each
   axis; equals the decision trees of reference src/locate_patch_low_R.cpp,
   src/locate_patch_high_R.cpp sampled as in src/t_c_g0_n.cpp:53-64 */
void oracle_patch_grid(int *grid_lo, int *grid_hi)
{
   for (int p = 0; p < TCG_NPATCH; p++) {
      const unsigned char *b = tcg_patch_box[p];
      int *g = p < TCG_NPATCH_LOW ? grid_lo : grid_hi;
      for (int i1 = b[0]; i1 < b[1]; i1++)
         for (int i2 = b[2]; i2 < b[3]; i2++) g[i1 * 128 + i2] = p;
   }
}

/* bias and width of each patch, TG = (2X - B) W (the reference's
   src/BW_by_patch.h holds the same numbers for the same boxes) */
void oracle_patch_bw(double *bw)
{
   for (int p = 0; p < TCG_NPATCH; p++) {
      const unsigned char *b = tcg_patch_box[p];
      bw[4 * p + 0] = (double)(b[0] + b[1]) / 128.0;
      bw[4 * p + 1] = 128.0 / (double)(b[1] - b[0]);
      bw[4 * p + 2] = (double)(b[2] + b[3]) / 128.0;
      bw[4 * p + 3] = 128.0 / (double)(b[3] - b[2]);
   }
}

/* the cartesian -> spherical coefficients the oracle applies (l = 0..3) */
void oracle_c2s(double *out)
{
   static const double v[GINT_C2S_SIZE] = {GINT_C2S_VALUES};
   for (int i = 0; i < GINT_C2S_SIZE; i++) out[i] = v[i];
}

/* ------------------------------------------------------------------ */
/* engine state                                                        */
/* ------------------------------------------------------------------ */

typedef struct {
   int defined;
   double R[3];
   int np;
   double Z[16];
   int nsub;
   int l[8], nl[8];
   double *K[8]; /* nl x np normalised coefficients */
} OSet;

typedef struct {
   int i, j, k, l, n1, n2;
   int sa, sb, sc, sd;
   int nprm;
   uint32_t *prm;
   unsigned Oq;
} ORow;

typedef struct {
   int code;
   int ld[4];
   unsigned off[4];
   int T[4];
} OKs;

typedef struct {
   int used;
   ORow *rows;
   int nrows, caprows;
   unsigned nQ;
   OKs *ks;
   int nks, capks;
   int out_mode, nl[4]; /* oracle_add_tra: integrals to out[dest ...] */
   uint64_t dest;
} OClass;

/* one class of a closed set quartet waiting to be integrated: its rows (it
   owns their primitive lists) and its K entries */
typedef struct {
   int c;
   ORow *rows;
   int nrows;
   unsigned nQ;
   OKs *ks;
   int nks;
   int out_mode, nl[4];
   uint64_t dest;
} OPendClass;

typedef struct {
   OPendClass *cls;
   int ncls;
} OPendSet;

/* per-thread counters of the integrated slots */
typedef struct {
   uint64_t n_slots, n_surv, n_prmq, checksum;
} OStats;

/* closed set quartets are integrated in batches of this many, one OpenMP
   thread per set quartet (round robin), each into its own K */
#define OPEND_BATCH 2048

typedef struct {
   int truncated;
   double R_cut;
   double *C0;
   int ld_C0, nder;
   int grid_lo[128 * 128], grid_hi[128 * 128];
   double bw[TCG_NPATCH * 4];
   double hf_fac;
   double cell[18];
   double *neighs;
   int nneighs;
   OSet *sets;
   int nsets;
   int nspin;
   size_t FP_size;
   double *P[2], *K[2];
   uint32_t *prm_tmp;
   int nprm_tmp, capprm;
   int cur[4], have_cur; /* sets of the open set quartet (rows with primitives) */
   OClass cls[NCLS];
   uint64_t n_slots, n_surv, n_prmq, n_qrtt, checksum;
   OPendSet *pend;
   int npend, cappend;
   double *out; /* oracle_set_output */
   uint64_t out_n;
} Oracle;

static void o_flush(Oracle *o);

void *oracle_create_handle(void)
{
   Oracle *o = (Oracle *)calloc(1, sizeof(Oracle));
   o->hf_fac = 1.0;
   for (int i = 0; i < 9; i++) o->cell[i] = o->cell[9 + i] = (i % 4 == 0) ? 1.0 : 0.0;
   oracle_patch_grid(o->grid_lo, o->grid_hi);
   oracle_patch_bw(o->bw);
   return o;
}

static void free_class(OClass *c)
{
   for (int r = 0; r < c->nrows; r++) free(c->rows[r].prm);
   c->nrows = 0;
   c->nks = 0;
   c->nQ = 0;
   c->used = 0;
}

static void free_pending(Oracle *o)
{
   for (int i = 0; i < o->npend; i++) {
      OPendSet *ps = &o->pend[i];
      for (int c = 0; c < ps->ncls; c++) {
         for (int r = 0; r < ps->cls[c].nrows; r++) free(ps->cls[c].rows[r].prm);
         free(ps->cls[c].rows);
         free(ps->cls[c].ks);
      }
      free(ps->cls);
   }
   o->npend = 0;
}

void oracle_destroy_handle(void *h)
{
   Oracle *o = (Oracle *)h;
   if (!o) return;
   free_pending(o);
   free(o->pend);
   for (int c = 0; c < NCLS; c++) {
      free_class(&o->cls[c]);
      free(o->cls[c].rows);
      free(o->cls[c].ks);
   }
   for (int s = 0; s < o->nsets; s++)
      for (int u = 0; u < o->sets[s].nsub; u++) free(o->sets[s].K[u]);
   free(o->sets);
   free(o->neighs);
   free(o->C0);
   free(o->P[0]); free(o->P[1]); free(o->K[0]); free(o->K[1]);
   free(o->prm_tmp);
   free(o);
}

void oracle_init(void *h) { (void)h; }

/* reference src/libGint.cpp:62-100 (the table is copied) */
void oracle_set_Potential_Truncated(void *h, double R_cut, double *C0, int ld_C0, int C0_size)
{
   Oracle *o = (Oracle *)h;
   o->truncated = 1;
   o->R_cut = R_cut;
   o->ld_C0 = ld_C0;
   o->nder = ld_C0 / TCG_NCOEF - 1;
   free(o->C0);
   o->C0 = (double *)malloc(sizeof(double) * (size_t)C0_size);
   memcpy(o->C0, C0, sizeof(double) * (size_t)C0_size);
}

void oracle_set_hf_fac(void *h, double fac) { ((Oracle *)h)->hf_fac = fac; }
void oracle_set_max_mem(void *h, int m) { (void)h; (void)m; }

void oracle_set_cell(void *h, bool periodic, double *cell_h, double *cell_i)
{
   Oracle *o = (Oracle *)h;
   (void)periodic;
   memcpy(o->cell, cell_h, 9 * sizeof(double));
   memcpy(o->cell + 9, cell_i, 9 * sizeof(double));
}

void oracle_set_neighs(void *h, double *neighs, int n)
{
   Oracle *o = (Oracle *)h;
   free(o->neighs);
   o->neighs = (double *)malloc(sizeof(double) * 3 * (size_t)n);
   memcpy(o->neighs, neighs, sizeof(double) * 3 * (size_t)n);
   o->nneighs = n;
}

static OSet *get_set(Oracle *o, int i)
{
   if (i >= o->nsets) {
      o->sets = (OSet *)realloc(o->sets, sizeof(OSet) * (size_t)(i + 1));
      memset(o->sets + o->nsets, 0, sizeof(OSet) * (size_t)(i + 1 - o->nsets));
      o->nsets = i + 1;
   }
   return &o->sets[i];
}

/* reference src/libGint.cpp:117-138 */
void oracle_set_Atom(void *h, int i, double *R, double *Z, int np)
{
   Oracle *o = (Oracle *)h;
   OSet *s = get_set(o, i);
   for (int u = 0; u < s->nsub; u++) free(s->K[u]);
   s->nsub = 0;
   s->defined = 1;
   memcpy(s->R, R, 3 * sizeof(double));
   s->np = np;
   memcpy(s->Z, Z, (size_t)np * sizeof(double));
}

/* reference src/extern_functions.cpp:95-110 */
double oracle_norm_psi(const double *K, const double *z, int np, int l)
{
   double ans = 0.0;
   for (int i = 0; i < np; i++)
      for (int j = 0; j < np; j++) ans += K[i] * K[j] / pow(z[i] + z[j], 1.5 + l);
   ans *= pow(M_PI, 1.5) / pow(2.0, l);
   return sqrt(ans);
}

/* reference src/extern_functions.cpp:114-136 and src/libGint.cpp:140-148 */
void oracle_set_Atom_L(void *h, int i, int l, int nl, double *K)
{
   Oracle *o = (Oracle *)h;
   OSet *s = get_set(o, i);
   int np = s->np;
   double *Kn = (double *)malloc(sizeof(double) * (size_t)(nl * np));
   for (int inl = 0; inl < nl; inl++) {
      double norm = oracle_norm_psi(&K[inl * np], s->Z, np, l);
      for (int ip = 0; ip < np; ip++) Kn[inl * np + ip] = K[inl * np + ip] / norm;
   }
   s->l[s->nsub] = l;
   s->nl[s->nsub] = nl;
   s->K[s->nsub] = Kn;
   s->nsub++;
}

/* reference src/libGint.cpp:436-455,555-563 (set_P also zeroes K) */
void oracle_set_P(void *h, double *P, int P_size)
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   o->nspin = 1;
   o->FP_size = (size_t)P_size;
   for (int s = 0; s < 2; s++) { free(o->P[s]); free(o->K[s]); o->P[s] = o->K[s] = NULL; }
   o->P[0] = (double *)malloc(sizeof(double) * o->FP_size);
   memcpy(o->P[0], P, sizeof(double) * o->FP_size);
   o->K[0] = (double *)calloc(o->FP_size, sizeof(double));
}

void oracle_set_P_polarized(void *h, double *Pa, double *Pb, int P_size)
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   o->nspin = 2;
   o->FP_size = (size_t)P_size;
   for (int s = 0; s < 2; s++) { free(o->P[s]); free(o->K[s]); }
   o->P[0] = (double *)malloc(sizeof(double) * o->FP_size);
   o->P[1] = (double *)malloc(sizeof(double) * o->FP_size);
   memcpy(o->P[0], Pa, sizeof(double) * o->FP_size);
   memcpy(o->P[1], Pb, sizeof(double) * o->FP_size);
   o->K[0] = (double *)calloc(o->FP_size, sizeof(double));
   o->K[1] = (double *)calloc(o->FP_size, sizeof(double));
}

/* reference src/libGint.cpp:581-610: K starts from the given values */
void oracle_set_K(void *h, double *K, int K_size)
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   free(o->K[0]);
   o->FP_size = (size_t)K_size;
   o->K[0] = (double *)malloc(sizeof(double) * o->FP_size);
   memcpy(o->K[0], K, sizeof(double) * o->FP_size);
}

void oracle_set_K_polarized(void *h, double *Ka, double *Kb, int K_size)
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   o->FP_size = (size_t)K_size;
   for (int s = 0; s < 2; s++) free(o->K[s]);
   o->K[0] = (double *)malloc(sizeof(double) * o->FP_size);
   o->K[1] = (double *)malloc(sizeof(double) * o->FP_size);
   memcpy(o->K[0], Ka, sizeof(double) * o->FP_size);
   memcpy(o->K[1], Kb, sizeof(double) * o->FP_size);
}

void oracle_get_K(void *h, double *K)
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   memcpy(K, o->K[0], sizeof(double) * o->FP_size);
}

void oracle_get_K_polarized(void *h, double *Ka, double *Kb)
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   memcpy(Ka, o->K[0], sizeof(double) * o->FP_size);
   memcpy(Kb, o->K[1], sizeof(double) * o->FP_size);
}

/* reference src/libGint.cpp:151-155 */
void oracle_add_prm(void *h, int ipa, int ipb, int ipc, int ipd)
{
   Oracle *o = (Oracle *)h;
   if (o->nprm_tmp == o->capprm) {
      o->capprm = o->capprm ? 2 * o->capprm : 4096;
      o->prm_tmp = (uint32_t *)realloc(o->prm_tmp, sizeof(uint32_t) * (size_t)o->capprm);
   }
   o->prm_tmp[o->nprm_tmp++] = (uint32_t)ipa | ((uint32_t)ipb << 8) | ((uint32_t)ipc << 16) | ((uint32_t)ipd << 24);
}

static int cls_index(int la, int lb, int lc, int ld) { return ((la * 4 + lb) * 4 + lc) * 4 + ld; }

/* reference src/libGint.cpp:157-236: one row per angular-momentum
   combination of the four sets, sharing the primitive list; Oq is the
   quartet offset inside the current set (all (n1,n2) rows share it) */
void oracle_add_shell(void *h, int i, int j, int k, int l, int n1, int n2)
{
   Oracle *o = (Oracle *)h;
   if (o->nprm_tmp == 0) return;
   OSet *S[4] = {&o->sets[i], &o->sets[j], &o->sets[k], &o->sets[l]};
   o->cur[0] = i; o->cur[1] = j; o->cur[2] = k; o->cur[3] = l;
   o->have_cur = 1;
   for (int sa = 0; sa < S[0]->nsub; sa++)
      for (int sb = 0; sb < S[1]->nsub; sb++)
         for (int sc = 0; sc < S[2]->nsub; sc++)
            for (int sd = 0; sd < S[3]->nsub; sd++) {
               int c = cls_index(S[0]->l[sa], S[1]->l[sb], S[2]->l[sc], S[3]->l[sd]);
               OClass *C = &o->cls[c];
               if (C->nrows == C->caprows) {
                  C->caprows = C->caprows ? 2 * C->caprows : 16;
                  C->rows = (ORow *)realloc(C->rows, sizeof(ORow) * (size_t)C->caprows);
               }
               ORow *r = &C->rows[C->nrows++];
               r->i = i; r->j = j; r->k = k; r->l = l; r->n1 = n1; r->n2 = n2;
               r->sa = sa; r->sb = sb; r->sc = sc; r->sd = sd;
               r->nprm = o->nprm_tmp;
               r->prm = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)o->nprm_tmp);
               memcpy(r->prm, o->prm_tmp, sizeof(uint32_t) * (size_t)o->nprm_tmp);
               r->Oq = C->nQ;
               C->used = 1;
            }
   o->nprm_tmp = 0;
}

void oracle_add_cell(void *h) { (void)h; }

/* reference src/libGint.cpp:240-248 */
void oracle_add_qrt(void *h, int la, int lb, int lc, int ld, int nla, int nlb, int nlc, int nld)
{
   Oracle *o = (Oracle *)h;
   OClass *C = &o->cls[cls_index(la, lb, lc, ld)];
   C->nQ += (unsigned)(nla * nlb * nlc * nld);
   C->used = 1;
}

/* reference src/libGint.cpp:250-283 (symm_fac coding) */
void oracle_add_qrtt(void *h, double symm_fac, int la, int lb, int lc, int ld,
                     int inla, int inlb, int inlc, int inld, int ld_ac, int ld_ad, int ld_bc, int ld_bd,
                     unsigned off_ac, unsigned off_ad, unsigned off_bc, unsigned off_bd,
                     bool Tac, bool Tad, bool Tbc, bool Tbd)
{
   Oracle *o = (Oracle *)h;
   (void)inla; (void)inlb; (void)inlc; (void)inld;
   OClass *C = &o->cls[cls_index(la, lb, lc, ld)];
   if (C->nks == C->capks) {
      C->capks = C->capks ? 2 * C->capks : 64;
      C->ks = (OKs *)realloc(C->ks, sizeof(OKs) * (size_t)C->capks);
   }
   OKs *e = &C->ks[C->nks++];
   int code;
   if (symm_fac < 0.1) code = 4;
   else if (symm_fac < 0.3) code = 3;
   else if (symm_fac < 0.75) code = 2;
   else if (symm_fac < 1.5) code = 1;
   else code = 0;
   e->code = code;
   e->ld[0] = ld_ac; e->ld[1] = ld_ad; e->ld[2] = ld_bc; e->ld[3] = ld_bd;
   e->off[0] = off_ac; e->off[1] = off_ad; e->off[2] = off_bc; e->off[3] = off_bd;
   e->T[0] = Tac; e->T[1] = Tad; e->T[2] = Tbc; e->T[3] = Tbd;
   C->used = 1;
   o->n_qrtt++;
}

/* every add_qrt / add_qrtt call of the open set quartet for K blocks laid
   out per pair of sets (the product's libgint_amd_add_set_blocks): element
   (f,g) of the (x,y) pair block is at off + (T ? f ld + g : g ld + f), the
   functions of a set shell by shell, contraction by contraction, 2l+1 each */
void oracle_add_set_blocks(void *h, double symm_fac, int ld_ac, int ld_ad, int ld_bc, int ld_bd,
                           unsigned off_ac, unsigned off_ad, unsigned off_bc, unsigned off_bd,
                           bool Tac, bool Tad, bool Tbc, bool Tbd)
{
   Oracle *o = (Oracle *)h;
   if (!o->have_cur) return;
   const OSet *S[4] = {&o->sets[o->cur[0]], &o->sets[o->cur[1]], &o->sets[o->cur[2]], &o->sets[o->cur[3]]};
   int f0[4][8];
   for (int t = 0; t < 4; t++) {
      int f = 0;
      for (int u = 0; u < S[t]->nsub; u++) { f0[t][u] = f; f += S[t]->nl[u] * (2 * S[t]->l[u] + 1); }
   }
   const unsigned base[4] = {off_ac, off_ad, off_bc, off_bd};
   const int ld[4] = {ld_ac, ld_ad, ld_bc, ld_bd};
   const bool T[4] = {Tac, Tad, Tbc, Tbd};
   for (int sa = 0; sa < S[0]->nsub; sa++)
      for (int sb = 0; sb < S[1]->nsub; sb++)
         for (int sc = 0; sc < S[2]->nsub; sc++)
            for (int sd = 0; sd < S[3]->nsub; sd++) {
               const int la = S[0]->l[sa], lb = S[1]->l[sb], lc = S[2]->l[sc], ld_ = S[3]->l[sd];
               const int nla = S[0]->nl[sa], nlb = S[1]->nl[sb], nlc = S[2]->nl[sc], nld = S[3]->nl[sd];
               oracle_add_qrt(h, la, lb, lc, ld_, nla, nlb, nlc, nld);
               for (int a = 0; a < nla; a++)
                  for (int b = 0; b < nlb; b++)
                     for (int c = 0; c < nlc; c++)
                        for (int d = 0; d < nld; d++) {
                           const long fa = f0[0][sa] + a * (2 * la + 1), fb = f0[1][sb] + b * (2 * lb + 1);
                           const long fc = f0[2][sc] + c * (2 * lc + 1), fd = f0[3][sd] + d * (2 * ld_ + 1);
                           const long rr[4] = {fa, fa, fb, fb}, cc[4] = {fc, fd, fc, fd};
                           unsigned off[4];
                           for (int q = 0; q < 4; q++)
                              off[q] = base[q] + (unsigned)(T[q] ? rr[q] * ld[q] + cc[q] : cc[q] * ld[q] + rr[q]);
                           oracle_add_qrtt(h, symm_fac, la, lb, lc, ld_, a, b, c, d, ld[0], ld[1], ld[2], ld[3],
                                           off[0], off[1], off[2], off[3], T[0], T[1], T[2], T[3]);
                        }
            }
}

/* integral output in the reference's TRA layout (the product's
   libgint_amd_set_output / add_tra / get_output) */
void oracle_set_output(void *h, double *out, uint64_t n)
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   o->out = out;
   o->out_n = n;
   if (out) memset(out, 0, sizeof(double) * (size_t)n);
}

void oracle_add_tra(void *h, int la, int lb, int lc, int ld, uint64_t dest)
{
   Oracle *o = (Oracle *)h;
   if (!o->have_cur) return;
   const int want[4] = {la, lb, lc, ld};
   OClass *C = &o->cls[cls_index(la, lb, lc, ld)];
   for (int t = 0; t < 4; t++) {
      const OSet *S = &o->sets[o->cur[t]];
      C->nl[t] = 0;
      for (int u = 0; u < S->nsub; u++)
         if (S->l[u] == want[t]) C->nl[t] = S->nl[u];
   }
   C->nQ = (unsigned)(C->nl[0] * C->nl[1] * C->nl[2] * C->nl[3]);
   C->out_mode = 1;
   C->dest = dest;
   C->used = 1;
}

void oracle_get_output(void *h) { o_flush((Oracle *)h); }

/* ------------------------------------------------------------------ */
/* integrals                                                           */
/* ------------------------------------------------------------------ */

/* cartesian component tables: index within shell L of (lx,ly,lz) is
   (L-lx)(L-lx+1)/2 + lz (descending x, then descending y) */
static int g_cart[MAXLL + 2][NCO(MAXLL + 1)][3];
static int g_idx[MAXLL + 2][MAXLL + 2][MAXLL + 2];
static int g_tables_done = 0;
static void cart_tables(void)
{
   if (g_tables_done) return;
   for (int L = 0; L <= MAXLL + 1; L++) {
      int i = 0;
      for (int lx = L; lx >= 0; lx--)
         for (int ly = L - lx; ly >= 0; ly--) {
            int lz = L - lx - ly;
            g_cart[L][i][0] = lx; g_cart[L][i][1] = ly; g_cart[L][i][2] = lz;
            g_idx[lx][ly][lz] = i;
            i++;
         }
   }
   g_tables_done = 1;
}
static int cidx(const int v[3]) { return g_idx[v[0]][v[1]][v[2]]; }
static int first_dir(const int v[3]) { return v[0] > 0 ? 0 : (v[1] > 0 ? 1 : 2); }

/* survivor checksum key of a (slot, class): a hash of the slot (set
   indices, images, primitive indices, cell) times an odd multiplier of the
   class, summed mod 2^64 over the surviving pairs (order-independent) */
#define MIX(z) do { z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL; z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL; z = z ^ (z >> 31); } while (0)
uint64_t oracle_slot_key(int i, int j, int k, int l, int n1, int n2,
                         int la, int lb, int lc, int ld,
                         int ipa, int ipb, int ipc, int ipd, int n3)
{
   uint64_t w0 = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
   uint64_t w1 = ((uint64_t)(uint32_t)k << 32) | (uint32_t)l;
   uint64_t w2 = ((uint64_t)n1 << 24) | ((uint64_t)n2 << 16);
   uint64_t w3 = ((uint64_t)((ipa << 12) | (ipb << 8) | (ipc << 4) | ipd) << 8) | (uint64_t)n3;
   uint64_t z = w0 + 0x9E3779B97F4A7C15ULL;
   MIX(z); z ^= w1; MIX(z); z ^= w2; MIX(z); z ^= w3; MIX(z);
   int code = ((la * 4 + lb) * 4 + lc) * 4 + ld;
   uint64_t m = 0xC1A55E5ULL + (uint64_t)code * 0x9E3779B97F4A7C15ULL;
   MIX(m);
   return z * (m | 1ULL);
}
#undef MIX

/* G_m(R,T), reference src/t_c_g0_n.cpp:67-180, returns branch */
static int o_tcg(const Oracle *o, double *res, double R, double T, int Nder, double Kfac)
{
   double upper = R * R + 11.0 * R + 50.0;
   double lower = R * R - 11.0 * R + 0.0;
   double X1, X2;
   int patch;
   if (T > upper) {
      for (int n = 0; n <= Nder; n++) res[n] = 0.0;
      return 0;
   }
   if (R <= 11.0) {
      X2 = R / 11.0;
      upper = R * R + 11.0 * R + 50.0;
      lower = 0.0;
      X1 = (T - lower) / (upper - lower);
      int i1 = (int)(X1 * 128);
      int i2 = (int)(X2 * 128);
      if (i1 == 128) i1 = 127;
      if (i2 == 128) i2 = 127;
      patch = o->grid_lo[i1 * 128 + i2];
   } else {
      if (T < lower) {
         for (int n = 0; n <= Nder; n++) res[n] = 0.0;
         return 1;
      }
      X2 = 11.0 / R;
      X1 = (T - lower) / (upper - lower);
      int i1 = (int)(X1 * 128);
      int i2 = (int)(X2 * 128);
      if (i1 == 128) i1 = 127;
      if (i2 == 128) i2 = 127;
      patch = o->grid_hi[i1 * 128 + i2];
   }
   const double B1 = o->bw[patch * 4 + 0], W1 = o->bw[patch * 4 + 1];
   const double B2 = o->bw[patch * 4 + 2], W2 = o->bw[patch * 4 + 3];
   const double *C0_row = &o->C0[(size_t)o->ld_C0 * patch];
   double TG1 = (2. * X1 - B1) * W1;
   double TG2 = (2. * X2 - B2) * W2;
   const double SQRT2 = 1.4142135623730950488;
   double T1[14], T2[14];
   T1[0] = 1.0; T2[0] = 1.0;
   T1[1] = SQRT2 * TG1; T2[1] = SQRT2 * TG2;
   T1[2] = 2. * TG1 * T1[1] - SQRT2;
   T2[2] = 2. * TG2 * T2[1] - SQRT2;
   for (int i = 3; i < 14; i++) {
      T1[i] = 2. * TG1 * T1[i - 1] - T1[i - 2];
      T2[i] = 2. * TG2 * T2[i - 1] - T2[i - 2];
   }
   for (int k = 0; k <= Nder; k++) {
      double tmp = 0.0;
      int jl = 0;
      for (int j = 0; j < 14; j++) {
         double dot = 0.0;
         for (int l = 0; l < 14 - j; l++) {
            dot += T1[l] * C0_row[k * 105 + jl];
            jl++;
         }
         tmp += dot * T2[j];
      }
      res[k] = tmp * Kfac;
   }
   return 2;
}

int oracle_tcg(void *h, double R, double T, int nder, double *G)
{
   Oracle *o = (Oracle *)h;
   int b = o_tcg(o, G, R, T, nder, 1.0);
   if (b == 1) o_fgamma0(nder, T, G, s_ftable(), 1.0);
   return b;
}

/* per-row geometry and contraction data */
typedef struct {
   const OSet *S[4];
   int L[4];
   int nl[4];
   const double *K[4];
   double A[3], B[3], C[3], D[3];  /* images as used by the HRR (C unshifted) */
   double ABs[3], CDs[3];
   const double *g1, *g2;
} RowGeo;

/* VRR workspace: V[La][Lc][m][ia*NCO(Lc)+ic] */
typedef struct {
   int LA, LC, L;
   size_t off[MAXLL + 1][MAXLL + 1];
   int mstride[MAXLL + 1][MAXLL + 1];
   double *buf;
} VWork;

static void vwork_init(VWork *w, int LA, int LC)
{
   w->LA = LA; w->LC = LC; w->L = LA + LC;
   size_t tot = 0;
   for (int a = 0; a <= LA; a++)
      for (int c = 0; c <= LC; c++) {
         w->off[a][c] = tot;
         w->mstride[a][c] = NCO(a) * NCO(c);
         tot += (size_t)(w->L - a - c + 1) * (size_t)(NCO(a) * NCO(c));
      }
   w->buf = (double *)malloc(sizeof(double) * tot);
}

#define VAT(w, a, c, m, ia, ic) ((w)->buf[(w)->off[a][c] + (size_t)(m) * (w)->mstride[a][c] + (size_t)(ia) * NCO(c) + (ic)])

/* Obara-Saika VRR for one slot, reference src/compute_VRR.cpp:111-345.
   pq: PA[3] WP[3] QC[3] WQ[3] i2ab mrab i2cd mrcd i2z; Fm[0..L] */
static void o_vrr(VWork *w, const double *Fm, const double *pq)
{
   const double *PA = pq, *WP = pq + 3, *QC = pq + 6, *WQ = pq + 9;
   const double i2ab = pq[12], mrab = pq[13], i2cd = pq[14], mrcd = pq[15], i2z = pq[16];
   const int L = w->L;
   for (int m = 0; m <= L; m++) VAT(w, 0, 0, m, 0, 0) = Fm[m];
   /* bra: (a+1|0) */
   for (int a = 1; a <= w->LA; a++)
      for (int m = 0; m <= L - a; m++)
         for (int ia = 0; ia < NCO(a); ia++) {
            const int *v = g_cart[a][ia];
            int d = first_dir(v);
            int v1[3] = {v[0], v[1], v[2]};
            v1[d]--;
            int i1 = cidx(v1);
            double val = PA[d] * VAT(w, a - 1, 0, m, i1, 0) + WP[d] * VAT(w, a - 1, 0, m + 1, i1, 0);
            int e2 = v1[d];
            if (e2 > 0) {
               int v2[3] = {v1[0], v1[1], v1[2]};
               v2[d]--;
               int i2 = cidx(v2);
               val += e2 * (i2ab * VAT(w, a - 2, 0, m, i2, 0) + mrab * VAT(w, a - 2, 0, m + 1, i2, 0));
            }
            VAT(w, a, 0, m, ia, 0) = val;
         }
   /* ket: (a|c+1) with the cross term */
   for (int c = 1; c <= w->LC; c++)
      for (int a = 0; a <= w->LA; a++)
         for (int m = 0; m <= L - a - c; m++)
            for (int ia = 0; ia < NCO(a); ia++)
               for (int ic = 0; ic < NCO(c); ic++) {
                  const int *va = g_cart[a][ia];
                  const int *vc = g_cart[c][ic];
                  int d = first_dir(vc);
                  int c1[3] = {vc[0], vc[1], vc[2]};
                  c1[d]--;
                  int j1 = cidx(c1);
                  double val = QC[d] * VAT(w, a, c - 1, m, ia, j1) + WQ[d] * VAT(w, a, c - 1, m + 1, ia, j1);
                  int f2 = c1[d];
                  if (f2 > 0) {
                     int c2[3] = {c1[0], c1[1], c1[2]};
                     c2[d]--;
                     int j2 = cidx(c2);
                     val += f2 * (i2cd * VAT(w, a, c - 2, m, ia, j2) + mrcd * VAT(w, a, c - 2, m + 1, ia, j2));
                  }
                  int e2 = va[d];
                  if (e2 > 0) {
                     int a1[3] = {va[0], va[1], va[2]};
                     a1[d]--;
                     val += e2 * i2z * VAT(w, a - 1, c - 1, m + 1, cidx(a1), j1);
                  }
                  VAT(w, a, c, m, ia, ic) = val;
               }
}

/* contracted cartesian integrals of one row: out[nlabcd][Nc] with layout
   [a][b][c][d]; reference src/compute_Fm.cpp:20-326 (prepare, Fm, Vm),
   src/compute_VRR.cpp:349-458 (VRR and the Fm[0] > 1e-12 gate),
   src/compute_ECO.cpp:216-272 (sum over n3), src/compute_ECO.cpp:12-214
   (contraction), src/compute_HRR.cpp:7-225 (HRR) */
static void o_row(const Oracle *o, OStats *st, const ORow *r, int la, int lb, int lc, int ld, double *out)
{
   cart_tables();
   RowGeo G;
   const int sidx[4] = {r->i, r->j, r->k, r->l};
   const int ssub[4] = {r->sa, r->sb, r->sc, r->sd};
   for (int t = 0; t < 4; t++) {
      G.S[t] = &o->sets[sidx[t]];
      G.L[t] = G.S[t]->l[ssub[t]];
      G.nl[t] = G.S[t]->nl[ssub[t]];
      G.K[t] = G.S[t]->K[ssub[t]];
   }
   const int Lsum = la + lb + lc + ld;
   const int Ng = o->nneighs;
   const double *neighs = o->neighs;
   const double *Ao = G.S[0]->R, *Bo = G.S[1]->R, *Co = G.S[2]->R, *Do = G.S[3]->R;
   const int nla = G.nl[0], nlb = G.nl[1], nlc = G.nl[2], nld = G.nl[3];
   const int nlabcd = nla * nlb * nlc * nld;
   const int npa = G.S[0]->np, npb = G.S[1]->np, npc = G.S[2]->np, npd = G.S[3]->np;

   /* CP2S source blocks (e0|f0), e in [la,la+lb], f in [lc,lc+ld] */
   int cpo[MAXLL + 1][MAXLL + 1];
   int ncp = 0;
   for (int e = la; e <= la + lb; e++)
      for (int f = lc; f <= lc + ld; f++) { cpo[e][f] = ncp; ncp += NCO(e) * NCO(f); }

   double *ABCDe = (double *)calloc((size_t)nlabcd * ncp, sizeof(double));
   double *AC0 = (double *)malloc(sizeof(double) * (size_t)ncp);
   double *ACr = (double *)malloc(sizeof(double) * (size_t)ncp);
   double *blockK = (double *)malloc(sizeof(double) * (size_t)nlabcd * ncp);
   VWork w;
   vwork_init(&w, la + lb, lc + ld);
   double Fm[MAXLL + 1 + 1];

   /* fixed geometry of the row (reference src/compute_Fm.cpp:79-95) */
   double ABs[3], CDs[3], A[3], B[3], C[3], D[3];
   o_pbc(Ao, Bo, o->cell, ABs);
   A[0] = Ao[0]; A[1] = Ao[1]; A[2] = Ao[2];
   B[0] = Ao[0] + ABs[0] + neighs[r->n1 * 3 + 0];
   B[1] = Ao[1] + ABs[1] + neighs[r->n1 * 3 + 1];
   B[2] = Ao[2] + ABs[2] + neighs[r->n1 * 3 + 2];
   o_pbc(Co, Do, o->cell, CDs);
   C[0] = Co[0]; C[1] = Co[1]; C[2] = Co[2];
   D[0] = Co[0] + CDs[0] + neighs[r->n2 * 3 + 0];
   D[1] = Co[1] + CDs[1] + neighs[r->n2 * 3 + 1];
   D[2] = Co[2] + CDs[2] + neighs[r->n2 * 3 + 2];

   int nblk = 0; /* primitives accumulated in blockK, flushed every 8 */
   memset(blockK, 0, sizeof(double) * (size_t)nlabcd * ncp);

   for (int p = 0; p < r->nprm; p++) {
      const uint32_t pk = r->prm[p];
      const int ipa = pk & 255, ipb = (pk >> 8) & 255, ipc = (pk >> 16) & 255, ipd = (pk >> 24) & 255;
      const double za = G.S[0]->Z[ipa], zb = G.S[1]->Z[ipb], zc = G.S[2]->Z[ipc], zd = G.S[3]->Z[ipd];
      /* prepare, reference src/compute_Fm.cpp:66-159 */
      const double zab = za + zb;
      const double inv_zab = 1. / zab;
      const double zcd = zc + zd;
      const double inv_zcd = 1. / zcd;
      const double z = zab + zcd;
      const double inv_z = 1. / z;
      double P[3], Q[3], PQs[3], shift[3], Cs[3], Ds[3];
      P[0] = (za * A[0] + zb * B[0]) * inv_zab;
      P[1] = (za * A[1] + zb * B[1]) * inv_zab;
      P[2] = (za * A[2] + zb * B[2]) * inv_zab;
      Q[0] = (zc * C[0] + zd * D[0]) * inv_zcd;
      Q[1] = (zc * C[1] + zd * D[1]) * inv_zcd;
      Q[2] = (zc * C[2] + zd * D[2]) * inv_zcd;
      o_pbc(Q, P, o->cell, PQs);
      shift[0] = P[0] - Q[0] - PQs[0];
      shift[1] = P[1] - Q[1] - PQs[1];
      shift[2] = P[2] - Q[2] - PQs[2];
      double normf = 34.986836655249725 * inv_zab * inv_zcd;
      double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
      double RAB2 = dx * dx + dy * dy + dz * dz;
      double expon_ab = -(za * zb) / (za + zb) * RAB2;
      dx = C[0] - D[0]; dy = C[1] - D[1]; dz = C[2] - D[2];
      double RCD2 = dx * dx + dy * dy + dz * dz;
      double expon_cd = -(zc * zd) / (zc + zd) * RCD2;
      double Zn = 1. / sqrt(z);
      double Kfactor = Zn * normf * oracle_exp(expon_ab + expon_cd);
      Cs[0] = Co[0] + shift[0];
      Cs[1] = Co[1] + shift[1];
      Cs[2] = Co[2] + shift[2];
      Ds[0] = Co[0] + CDs[0] + neighs[r->n2 * 3 + 0] + shift[0];
      Ds[1] = Co[1] + CDs[1] + neighs[r->n2 * 3 + 1] + shift[1];
      Ds[2] = Co[2] + CDs[2] + neighs[r->n2 * 3 + 2] + shift[2];
      Q[0] = (zc * Cs[0] + zd * Ds[0]) * inv_zcd;
      Q[1] = (zc * Cs[1] + zd * Ds[1]) * inv_zcd;
      Q[2] = (zc * Cs[2] + zd * Ds[2]) * inv_zcd;
      double PA[3] = {P[0] - A[0], P[1] - A[1], P[2] - A[2]};
      double QC[3] = {Q[0] - Cs[0], Q[1] - Cs[1], Q[2] - Cs[2]};
      /* compute_Fm / compute_Vm scalars, reference src/compute_Fm.cpp:202-210,281-290 */
      const double zcd2 = zc + zd;
      const double rho = zab * zcd2 * inv_z;
      const double Rr = o->R_cut * sqrt(rho);
      const double izab = 1. / zab, i2ab = 0.5 * izab, mrab = -1. * rho * izab * i2ab;
      const double izcd = 1. / zcd2, i2cd = 0.5 * izcd, mrcd = -1. * rho * izcd * i2cd;
      const double i2z = 0.5 * inv_z;
      st->n_prmq++;

      memset(AC0, 0, sizeof(double) * (size_t)ncp);
      memset(ACr, 0, sizeof(double) * (size_t)ncp);
      for (int n3 = 0; n3 < Ng; n3++) {
         double PQ[3];
         PQ[0] = P[0] - Q[0] - neighs[n3 * 3 + 0];
         PQ[1] = P[1] - Q[1] - neighs[n3 * 3 + 1];
         PQ[2] = P[2] - Q[2] - neighs[n3 * 3 + 2];
         double rpq2 = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
         double T = rho * rpq2;
         if (o->truncated) {
            int br = o_tcg(o, Fm, Rr, T, Lsum, Kfactor);
            if (br == 1) o_fgamma0(Lsum, T, Fm, s_ftable(), Kfactor);
         } else {
            o_fgamma0(Lsum, T, Fm, s_ftable(), Kfactor);
         }
         st->n_slots++;
         double *dst = (n3 == 0) ? AC0 : ACr;
         if (Fm[0] > 1.e-12) {
            st->n_surv++;
            st->checksum += oracle_slot_key(r->i, r->j, r->k, r->l, r->n1, r->n2, la, lb, lc, ld, ipa, ipb, ipc, ipd, n3);
            double Qp[3] = {Q[0] + neighs[n3 * 3 + 0], Q[1] + neighs[n3 * 3 + 1], Q[2] + neighs[n3 * 3 + 2]};
            double W[3];
            W[0] = (zab * P[0] + zcd2 * Qp[0]) * inv_z;
            W[1] = (zab * P[1] + zcd2 * Qp[1]) * inv_z;
            W[2] = (zab * P[2] + zcd2 * Qp[2]) * inv_z;
            double pq[17] = {PA[0], PA[1], PA[2], W[0] - P[0], W[1] - P[1], W[2] - P[2],
                             QC[0], QC[1], QC[2], W[0] - Qp[0], W[1] - Qp[1], W[2] - Qp[2],
                             i2ab, mrab, i2cd, mrcd, i2z};
            o_vrr(&w, Fm, pq);
            for (int e = la; e <= la + lb; e++)
               for (int f = lc; f <= lc + ld; f++) {
                  double *blk = dst + cpo[e][f];
                  for (int ie = 0; ie < NCO(e); ie++)
                     for (int jf = 0; jf < NCO(f); jf++) blk[ie * NCO(f) + jf] += VAT(&w, e, f, 0, ie, jf);
               }
         } else if (la == 0 && lc == 0) {
            /* a screened slot still carries its ssss(0) value,
               reference src/compute_VRR.cpp:384-385 */
            dst[cpo[0][0]] += Fm[0];
         }
      }
      for (int j = 0; j < ncp; j++) AC0[j] += ACr[j];

      /* contraction, K = Ka Kb Kc Kd, partial sums of 8 primitives */
      for (int lab = 0; lab < nlabcd; lab++) {
         int a = lab / (nlb * nlc * nld), b = (lab / (nlc * nld)) % nlb, c = (lab / nld) % nlc, d = lab % nld;
         double Kp = G.K[0][a * npa + ipa] * G.K[1][b * npb + ipb] * G.K[2][c * npc + ipc] * G.K[3][d * npd + ipd];
         double *dstk = blockK + (size_t)lab * ncp;
         for (int j = 0; j < ncp; j++) dstk[j] += Kp * AC0[j];
      }
      nblk++;
      if (nblk == 8 || p == r->nprm - 1) {
         for (size_t j = 0; j < (size_t)nlabcd * ncp; j++) ABCDe[j] += blockK[j];
         memset(blockK, 0, sizeof(double) * (size_t)nlabcd * ncp);
         nblk = 0;
      }
   }

   /* HRR per contraction combination, reference src/compute_HRR.cpp:7-99 */
   const double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
   const double CD[3] = {C[0] - D[0], C[1] - D[1], C[2] - D[2]};
   const int na = NCO(la), nb = NCO(lb), nc = NCO(lc), nd = NCO(ld);
   const int Nc = na * nb * nc * nd;
   /* ket step layers Y[d'][c'][ic][id], bra step layers W[b'][a'][ia][ib] */
   for (int lab = 0; lab < nlabcd; lab++) {
      const double *E = ABCDe + (size_t)lab * ncp;
      double *o_l = out + (size_t)lab * Nc;
      /* Z[e][ie][ic][id] after the ket HRR */
      int zoff[MAXLL + 1];
      int ztot = 0;
      for (int e = la; e <= la + lb; e++) { zoff[e] = ztot; ztot += NCO(e) * nc * nd; }
      double *Z = (double *)malloc(sizeof(double) * (size_t)ztot);
      for (int e = la; e <= la + lb; e++) {
         /* layer d'=0: blocks c' in [lc, lc+ld] */
         double *Yprev[MAXLL + 2] = {0}, *Ycur[MAXLL + 2] = {0};
         for (int cp = lc; cp <= lc + ld; cp++) {
            Yprev[cp] = (double *)malloc(sizeof(double) * (size_t)(NCO(e) * NCO(cp)));
            for (int ie = 0; ie < NCO(e); ie++)
               for (int ic = 0; ic < NCO(cp); ic++)
                  Yprev[cp][ie * NCO(cp) + ic] = E[cpo[e][cp] + ie * NCO(cp) + ic];
         }
         for (int dd = 1; dd <= ld; dd++) {
            for (int cp = lc; cp <= lc + ld - dd; cp++) {
               Ycur[cp] = (double *)malloc(sizeof(double) * (size_t)(NCO(e) * NCO(cp) * NCO(dd)));
               for (int ie = 0; ie < NCO(e); ie++)
                  for (int ic = 0; ic < NCO(cp); ic++)
                     for (int id = 0; id < NCO(dd); id++) {
                        const int *vd = g_cart[dd][id];
                        int dir = first_dir(vd);
                        int dm[3] = {vd[0], vd[1], vd[2]};
                        dm[dir]--;
                        int idm = cidx(dm);
                        int vcp[3] = {g_cart[cp][ic][0], g_cart[cp][ic][1], g_cart[cp][ic][2]};
                        vcp[dir]++;
                        int icp = cidx(vcp);
                        const int ndm = NCO(dd - 1);
                        double v1 = Yprev[cp + 1][(ie * NCO(cp + 1) + icp) * ndm + idm];
                        double v0 = Yprev[cp][(ie * NCO(cp) + ic) * ndm + idm];
                        Ycur[cp][(ie * NCO(cp) + ic) * NCO(dd) + id] = v1 + CD[dir] * v0;
                     }
            }
            for (int cp = lc; cp <= lc + ld; cp++) { free(Yprev[cp]); Yprev[cp] = Ycur[cp]; Ycur[cp] = NULL; }
         }
         /* Yprev[lc] is (e 0|lc ld) */
         for (int ie = 0; ie < NCO(e); ie++)
            for (int icd = 0; icd < nc * nd; icd++) Z[zoff[e] + ie * nc * nd + icd] = Yprev[lc][ie * nc * nd + icd];
         for (int cp = lc; cp <= lc + ld; cp++) free(Yprev[cp]);
      }
      /* bra HRR for every (ic,id) */
      for (int icd = 0; icd < nc * nd; icd++) {
         double *Wp[MAXLL + 2] = {0}, *Wc[MAXLL + 2] = {0};
         for (int ap = la; ap <= la + lb; ap++) {
            Wp[ap] = (double *)malloc(sizeof(double) * (size_t)NCO(ap));
            for (int ia = 0; ia < NCO(ap); ia++) Wp[ap][ia] = Z[zoff[ap] + ia * nc * nd + icd];
         }
         for (int bb = 1; bb <= lb; bb++) {
            for (int ap = la; ap <= la + lb - bb; ap++) {
               Wc[ap] = (double *)malloc(sizeof(double) * (size_t)(NCO(ap) * NCO(bb)));
               for (int ia = 0; ia < NCO(ap); ia++)
                  for (int ib = 0; ib < NCO(bb); ib++) {
                     const int *vb = g_cart[bb][ib];
                     int dir = first_dir(vb);
                     int bm[3] = {vb[0], vb[1], vb[2]};
                     bm[dir]--;
                     int ibm = cidx(bm);
                     int vap[3] = {g_cart[ap][ia][0], g_cart[ap][ia][1], g_cart[ap][ia][2]};
                     vap[dir]++;
                     int iap = cidx(vap);
                     const int nbm = NCO(bb - 1);
                     double v1 = Wp[ap + 1][iap * nbm + ibm];
                     double v0 = Wp[ap][ia * nbm + ibm];
                     Wc[ap][ia * NCO(bb) + ib] = v1 + AB[dir] * v0;
                  }
            }
            for (int ap = la; ap <= la + lb; ap++) { free(Wp[ap]); Wp[ap] = Wc[ap]; Wc[ap] = NULL; }
         }
         for (int ia = 0; ia < na; ia++)
            for (int ib = 0; ib < nb; ib++) o_l[(ia * nb + ib) * nc * nd + icd] = Wp[la][ia * nb + ib];
         for (int ap = la; ap <= la + lb; ap++) free(Wp[ap]);
      }
      free(Z);
   }
   free(w.buf);
   free(ABCDe);
   free(AC0);
   free(ACr);
   free(blockK);
}

/* cartesian -> spherical, four stages each rotating the first index to the
   last, reference src/compute_SPH.cpp:58-237 */
static void o_sph(int la, int lb, int lc, int ld, const double *cart, double *sph, double *tmp1, double *tmp2)
{
   const int nc_[4] = {NCO(la), NCO(lb), NCO(lc), NCO(ld)};
   const int ns_[4] = {2 * la + 1, 2 * lb + 1, 2 * lc + 1, 2 * ld + 1};
   const int ls[4] = {la, lb, lc, ld};
   const double *in = cart;
   /* dimensions of the current tensor, first index first */
   int dims[4] = {nc_[0], nc_[1], nc_[2], nc_[3]};
   for (int s = 0; s < 4; s++) {
      const double *c2s = gint_c2s + gint_c2s_off[ls[s]];
      const int k = dims[0];
      const int n = dims[1] * dims[2] * dims[3];
      const int m = ns_[s];
      double *o = (s == 3) ? sph : ((s % 2 == 0) ? tmp1 : tmp2);
      for (int j = 0; j < n; j++)
         for (int i = 0; i < m; i++) {
            double acc = 0.0;
            for (int q = 0; q < k; q++) acc += c2s[i * k + q] * in[q * n + j];
            o[j * m + i] = acc;
         }
      in = o;
      int nd[4] = {dims[1], dims[2], dims[3], m};
      memcpy(dims, nd, sizeof dims);
   }
}

static const double o_symm_factors[5] = {2.0, 1.0, 0.5, 0.25, 0.0};

/* exchange update, reference src/compute_KS.cpp:8-117 */
static void o_ks(const OKs *e, int la, int lb, int lc, int ld, const double *I, const double *P, double *K, double hf_fac)
{
   const int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1, nsd = 2 * ld + 1;
   const int nscd = nsc * nsd, nsbcd = nsb * nscd, nsabcd = nsa * nsbcd;
   const double fac = o_symm_factors[e->code] * hf_fac;
   const unsigned ldac = (unsigned)e->ld[0], ldad = (unsigned)e->ld[1], ldbc = (unsigned)e->ld[2], ldbd = (unsigned)e->ld[3];
   for (int t = 0; t < nsabcd; t++) {
      const unsigned isa = t / nsbcd, isb = (t / nscd) % nsb, isc = (t / nsd) % nsc, isd = t % nsd;
      const unsigned iac = !e->T[0] ? isc * ldac + isa : isa * ldac + isc;
      const unsigned iad = !e->T[1] ? isd * ldad + isa : isa * ldad + isd;
      const unsigned ibc = !e->T[2] ? isc * ldbc + isb : isb * ldbc + isc;
      const unsigned ibd = !e->T[3] ? isd * ldbd + isb : isb * ldbd + isd;
      const unsigned Iac = e->off[0] + iac, Iad = e->off[1] + iad, Ibc = e->off[2] + ibc, Ibd = e->off[3] + ibd;
      const double iabcd = -fac * I[t];
      const double kbd = iabcd * P[Iac];
      const double kbc = iabcd * P[Iad];
      const double kad = iabcd * P[Ibc];
      const double kac = iabcd * P[Ibd];
      K[Ibd] += kbd;
      K[Ibc] += kbc;
      K[Iad] += kad;
      K[Iac] += kac;
   }
}

/* the per-class chain of reference src/libGint.cpp:705-997 for one closed
   set quartet, into K (one thread's copy) */
static void o_integrate_set(const Oracle *o, const OPendSet *ps, double *K[2], OStats *st)
{
   for (int ci = 0; ci < ps->ncls; ci++) {
      const OPendClass *C = &ps->cls[ci];
      const int c = C->c;
      const int la = c / 64, lb = (c / 16) % 4, lc = (c / 4) % 4, ld = c % 4;
      const int Nc = NCO(la) * NCO(lb) * NCO(lc) * NCO(ld);
      const int Ns = (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1);
      if (C->nQ == 0 || !o->P[0]) continue;
      double *ABCD0 = (double *)calloc((size_t)C->nQ * Nc, sizeof(double));
      for (int r = 0; r < C->nrows; r++) {
         const ORow *row = &C->rows[r];
         const OSet *S[4] = {&o->sets[row->i], &o->sets[row->j], &o->sets[row->k], &o->sets[row->l]};
         int nlabcd = S[0]->nl[row->sa] * S[1]->nl[row->sb] * S[2]->nl[row->sc] * S[3]->nl[row->sd];
         if (row->Oq + (unsigned)nlabcd > C->nQ) continue; /* no quartet declared for it */
         double *out = (double *)malloc(sizeof(double) * (size_t)nlabcd * Nc);
         o_row(o, st, row, la, lb, lc, ld, out);
         for (size_t t = 0; t < (size_t)nlabcd * Nc; t++) ABCD0[(size_t)row->Oq * Nc + t] += out[t];
         free(out);
      }
      double *SPH = (double *)malloc(sizeof(double) * (size_t)C->nQ * Ns);
      double *t1 = (double *)malloc(sizeof(double) * (size_t)Nc * 4);
      double *t2 = (double *)malloc(sizeof(double) * (size_t)Nc * 4);
      for (unsigned q = 0; q < C->nQ; q++) o_sph(la, lb, lc, ld, ABCD0 + (size_t)q * Nc, SPH + (size_t)q * Ns, t1, t2);
      if (C->out_mode) {
         /* reference src/compute_TRA.cpp:36-47: SPHER [ila,ilb,ilc,ild][isa,isb,isc,isd]
            to OUT [ila,isa,ilb,isb,ilc,isc,ild,isd] */
         const int nla = C->nl[0], nlb = C->nl[1], nlc = C->nl[2], nld = C->nl[3];
         const int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1, nsd = 2 * ld + 1;
         for (int q = 0; q < (int)C->nQ; q++) {
            const int ia = q / (nlb * nlc * nld), ib = (q / (nlc * nld)) % nlb, ic = (q / nld) % nlc, id = q % nld;
            for (int t = 0; t < Ns; t++) {
               const int isa = t / (nsb * nsc * nsd), isb = (t / (nsc * nsd)) % nsb, isc = (t / nsd) % nsc, isd = t % nsd;
               const size_t j = ((((((size_t)ia * nsa + isa) * nlb + ib) * nsb + isb) * nlc + ic) * nsc + isc) *
                                   (size_t)(nld * nsd) + (size_t)id * nsd + isd;
               o->out[C->dest + j] = SPH[(size_t)q * Ns + t];
            }
         }
      }
      for (int e = 0; e < C->nks && (unsigned)e < C->nQ; e++)
         for (int s = 0; s < o->nspin; s++)
            o_ks(&C->ks[e], la, lb, lc, ld, SPH + (size_t)e * Ns, o->P[s], K[s], o->hf_fac);
      free(t1); free(t2); free(SPH); free(ABCD0);
   }
}

/* integrates the pending set quartets: set quartet i on thread i mod T, each
   thread into a private K, then the private K summed into K in thread order
   (the result depends on T only through the order of that sum) */
static void o_flush(Oracle *o)
{
   if (o->npend == 0) return;
   cart_tables();
   (void)s_ftable();
   const int T = omp_get_max_threads();
   const size_t n = o->FP_size;
   const int nspin = o->nspin > 0 ? o->nspin : 1;
   double *priv = NULL;
   if (o->P[0]) priv = (double *)calloc((size_t)T * (size_t)nspin * (n ? n : 1), sizeof(double));
   OStats *st = (OStats *)calloc((size_t)T, sizeof(OStats));
#pragma omp parallel num_threads(T)
   {
      const int t = omp_get_thread_num();
      double *Kt[2] = {NULL, NULL};
      if (priv)
         for (int s = 0; s < nspin; s++) Kt[s] = priv + ((size_t)t * nspin + s) * n;
#pragma omp for schedule(static, 1)
      for (int i = 0; i < o->npend; i++) o_integrate_set(o, &o->pend[i], Kt, &st[t]);
   }
   if (priv) {
      for (int s = 0; s < o->nspin; s++)
         for (int t = 0; t < T; t++) {
            const double *Kt = priv + ((size_t)t * nspin + s) * n;
            for (size_t x = 0; x < n; x++) o->K[s][x] += Kt[x];
         }
   }
   for (int t = 0; t < T; t++) {
      o->n_slots += st[t].n_slots;
      o->n_surv += st[t].n_surv;
      o->n_prmq += st[t].n_prmq;
      o->checksum += st[t].checksum;
   }
   free(st);
   free(priv);
   free_pending(o);
}

/* reference src/libGint.cpp:285-414: the set quartet is closed; its classes
   are integrated with the next batch */
void oracle_add_set(void *h)
{
   Oracle *o = (Oracle *)h;
   o->have_cur = 0;
   int nused = 0;
   for (int c = 0; c < NCLS; c++) nused += o->cls[c].used;
   if (nused == 0) return;
   if (o->npend == o->cappend) {
      o->cappend = o->cappend ? 2 * o->cappend : 64;
      o->pend = (OPendSet *)realloc(o->pend, sizeof(OPendSet) * (size_t)o->cappend);
   }
   OPendSet *ps = &o->pend[o->npend++];
   ps->cls = (OPendClass *)malloc(sizeof(OPendClass) * (size_t)nused);
   ps->ncls = 0;
   for (int c = 0; c < NCLS; c++) {
      OClass *C = &o->cls[c];
      if (!C->used) continue;
      OPendClass *pc = &ps->cls[ps->ncls++];
      pc->c = c;
      pc->nrows = C->nrows;
      pc->rows = (ORow *)malloc(sizeof(ORow) * (size_t)(C->nrows ? C->nrows : 1));
      memcpy(pc->rows, C->rows, sizeof(ORow) * (size_t)C->nrows);   /* the primitive lists move too */
      pc->nQ = C->nQ;
      pc->nks = C->nks;
      pc->ks = (OKs *)malloc(sizeof(OKs) * (size_t)(C->nks ? C->nks : 1));
      memcpy(pc->ks, C->ks, sizeof(OKs) * (size_t)C->nks);
      pc->out_mode = C->out_mode;
      memcpy(pc->nl, C->nl, sizeof pc->nl);
      pc->dest = C->dest;
      C->out_mode = 0;
      C->nrows = 0;
      C->nks = 0;
      C->nQ = 0;
      C->used = 0;
   }
   if (o->npend >= OPEND_BATCH) o_flush(o);
}

void oracle_dispatch(void *h) { o_flush((Oracle *)h); }

void oracle_stats(void *h, uint64_t out[5])
{
   Oracle *o = (Oracle *)h;
   o_flush(o);
   out[0] = o->n_slots;
   out[1] = o->n_surv;
   out[2] = o->n_prmq;
   out[3] = o->n_qrtt;
   out[4] = o->checksum;
}

/* test hook: contracted cartesian integrals [nlabcd][Nc] of one row of
   sets (i,j,k,l), subshells (sa..sd), images (n1,n2), primitive list */
void oracle_row_cart(void *h, int i, int j, int k, int l, int n1, int n2, int sa, int sb, int sc, int sd,
                     int nprm, const uint32_t *prm, double *out)
{
   Oracle *o = (Oracle *)h;
   ORow r;
   r.i = i; r.j = j; r.k = k; r.l = l; r.n1 = n1; r.n2 = n2;
   r.sa = sa; r.sb = sb; r.sc = sc; r.sd = sd;
   r.nprm = nprm;
   r.prm = (uint32_t *)prm;
   r.Oq = 0;
   OStats st = {0, 0, 0, 0};
   o_row(o, &st, &r, o->sets[i].l[sa], o->sets[j].l[sb], o->sets[k].l[sc], o->sets[l].l[sd], out);
}
