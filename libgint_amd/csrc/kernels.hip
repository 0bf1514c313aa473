// gfx950 kernels of the K-build.  A launch covers one chunk of one quartet
// type (the shell structure of the four sets of a set quartet).
//
//   k_fm_r         one wavefront per primitive quartet (two when the cells
//                  fit half the lanes).  Lanes evaluate
//                  G_m(R,T) (or F_m) of 64 lattice cells at once and decide
//                  Fm[0] > 1e-12 for every class L of the type; surviving
//                  cells run the McMurchie-Davidson Hermite recursion
//                  R^(n)_tuv in LDS and the n = 0 layer is summed over cells
//                  in registers.  One Hermite -> cartesian transform per
//                  primitive quartet then gives the cell-summed (e0|f0)
//                  blocks of every class of the type.  Replaces, for all
//                  classes at once, reference src/compute_Fm.cpp (prepare,
//                  Fm, Vm), src/compute_VRR.cpp and the cell sum of
//                  src/compute_ECO.cpp:216-272.
//   k_fm_r<true,.> slots whose screening differs between the classes of the
//                  type (Fm[0] within rounding of 1e-12 at different L): the
//                  slot's blocks alone, added by the consumers to the classes
//                  that keep it.
//   k_gqs          one workgroup per set quartet, its classes in turn, a
//                  batch of contractions at once: contraction over primitives
//                  (src/compute_ECO.cpp:12-214), HRR of each (n1,n2) row
//                  (src/compute_HRR.cpp) summed on chip, cartesian ->
//                  spherical (src/compute_SPH.cpp:58-237), and the exchange
//                  update of K (src/compute_KS.cpp:8-117) reduced over the
//                  set quartet's contractions before the atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "device.hpp"
#include "gint_math.h"
#include "c2s_table.h"

namespace gint {

// ---------------------------------------------------------------------------
// cartesian and Hermite index tables, built at compile time
// ---------------------------------------------------------------------------
constexpr int NSHELL = 13;  // shells 0..12

struct CartTables {
   unsigned char comp[NSHELL][91][3];  // lx ly lz of component i of shell L
   unsigned char first[NSHELL][91];    // first non-zero direction
   signed char minus[NSHELL][91][3];   // index of i - 1_d in shell L-1, -1 if none
   signed char plus[7][28][3];         // index of i + 1_d in shell L+1
};

constexpr int cart_index(int L, int lx, int lz) { return (L - lx) * (L - lx + 1) / 2 + lz; }

constexpr CartTables make_cart_tables()
{
   CartTables t{};
   for (int L = 0; L < NSHELL; L++) {
      int i = 0;
      for (int lx = L; lx >= 0; lx--)
         for (int ly = L - lx; ly >= 0; ly--) {
            int lz = L - lx - ly;
            int v[3] = {lx, ly, lz};
            for (int d = 0; d < 3; d++) t.comp[L][i][d] = (unsigned char)v[d];
            t.first[L][i] = (unsigned char)(lx > 0 ? 0 : (ly > 0 ? 1 : 2));
            for (int d = 0; d < 3; d++) {
               if (v[d] > 0) {
                  int w[3] = {v[0], v[1], v[2]};
                  w[d]--;
                  t.minus[L][i][d] = (signed char)cart_index(L - 1, w[0], w[2]);
               } else {
                  t.minus[L][i][d] = -1;
               }
               if (L < 7) {
                  int u[3] = {v[0], v[1], v[2]};
                  u[d]++;
                  t.plus[L][i][d] = (signed char)cart_index(L + 1, u[0], u[2]);
               }
            }
            i++;
         }
   }
   return t;
}

// Hermite triples (t,u,v), ordered by N = t+u+v, then like the cartesian
// components of shell N; the first nherm(K) entries are those with N <= K.
struct HermEntry {
   short m1, m2;          // index of tuv - 1_dir and tuv - 2_dir (-1 if none)
   unsigned char dir;     // recursion direction: first non-zero of t, u, v
   unsigned char coef;    // component along dir, minus one
   unsigned char N, pad;
};
struct HermTables {
   HermEntry e[MAXHERM];
   unsigned char tuv[MAXHERM][3];
   short idx[NSHELL][NSHELL][NSHELL];
};

constexpr HermTables make_herm_tables()
{
   HermTables h{};
   int n = 0;
   for (int N = 0; N < NSHELL; N++)
      for (int t = N; t >= 0; t--)
         for (int u = N - t; u >= 0; u--) {
            const int v = N - t - u;
            h.tuv[n][0] = (unsigned char)t;
            h.tuv[n][1] = (unsigned char)u;
            h.tuv[n][2] = (unsigned char)v;
            h.idx[t][u][v] = (short)n;
            n++;
         }
   for (int i = 0; i < n; i++) {
      const int t = h.tuv[i][0], u = h.tuv[i][1], v = h.tuv[i][2];
      HermEntry e{};
      e.N = (unsigned char)(t + u + v);
      e.m1 = e.m2 = -1;
      if (t + u + v > 0) {
         const int d = t > 0 ? 0 : (u > 0 ? 1 : 2);
         int w[3] = {t, u, v};
         const int c = w[d];
         e.dir = (unsigned char)d;
         e.coef = (unsigned char)(c - 1);
         w[d]--;
         e.m1 = h.idx[w[0]][w[1]][w[2]];
         if (c >= 2) {
            w[d]--;
            e.m2 = h.idx[w[0]][w[1]][w[2]];
         }
      }
      h.e[i] = e;
   }
   return h;
}

__constant__ CartTables c_cart = make_cart_tables();
__constant__ HermTables c_herm = make_herm_tables();
__constant__ double c_c2s[GINT_C2S_SIZE] = {GINT_C2S_VALUES};
__constant__ int c_c2s_off[4] = {GINT_C2S_OFF_0, GINT_C2S_OFF_1, GINT_C2S_OFF_2, GINT_C2S_OFF_3};

__device__ __forceinline__ int d_nco(int l) { return (l + 1) * (l + 2) / 2; }
__device__ __forceinline__ int d_nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// ---------------------------------------------------------------------------
// per-primitive quantities, in the order of reference
// src/compute_Fm.cpp:66-159 and 202-210, 281-290 (no contraction: these
// feed the Fm[0] > 1e-12 decision)
// ---------------------------------------------------------------------------
struct PrimGeo {
   double P[3], Q[3], PA[3], QC[3];
   double Kfac, zab, zcd, rho, Rr;
   double i2ab, i2cd;
};

__device__ void prim_prepare(const double *Ao, const double *Bo, const double *Co, const double *Do,
                             double za, double zb, double zc, double zd, int n1, int n2,
                             const DevTables &t, PrimGeo &g)
{
   GINT_NO_CONTRACT
   const double *neighs = t.neighs;
   const double zab = za + zb;
   const double inv_zab = 1. / zab;
   const double zcd = zc + zd;
   const double inv_zcd = 1. / zcd;
   const double z = zab + zcd;
   const double inv_z = 1. / z;
   double A[3], B[3], C[3], D[3], bimg[3], dimg[3], qimg[3], kshift[3], Cn[3], Dn[3], P[3], Q[3];
   gint_pbc(Ao, Bo, t.cell, bimg);
   A[0] = Ao[0]; A[1] = Ao[1]; A[2] = Ao[2];
   B[0] = Ao[0] + bimg[0] + neighs[n1 * 3 + 0];
   B[1] = Ao[1] + bimg[1] + neighs[n1 * 3 + 1];
   B[2] = Ao[2] + bimg[2] + neighs[n1 * 3 + 2];
   P[0] = (za * A[0] + zb * B[0]) * inv_zab;
   P[1] = (za * A[1] + zb * B[1]) * inv_zab;
   P[2] = (za * A[2] + zb * B[2]) * inv_zab;
   gint_pbc(Co, Do, t.cell, dimg);
   C[0] = Co[0]; C[1] = Co[1]; C[2] = Co[2];
   D[0] = Co[0] + dimg[0] + neighs[n2 * 3 + 0];
   D[1] = Co[1] + dimg[1] + neighs[n2 * 3 + 1];
   D[2] = Co[2] + dimg[2] + neighs[n2 * 3 + 2];
   Q[0] = (zc * C[0] + zd * D[0]) * inv_zcd;
   Q[1] = (zc * C[1] + zd * D[1]) * inv_zcd;
   Q[2] = (zc * C[2] + zd * D[2]) * inv_zcd;
   gint_pbc(Q, P, t.cell, qimg);
   kshift[0] = P[0] - Q[0] - qimg[0];
   kshift[1] = P[1] - Q[1] - qimg[1];
   kshift[2] = P[2] - Q[2] - qimg[2];
   const double pref = GINT_TWO_PI_52 * inv_zab * inv_zcd;
   double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
   const double ab2 = dx * dx + dy * dy + dz * dz;
   const double eab = -(za * zb) / (za + zb) * ab2;
   dx = C[0] - D[0]; dy = C[1] - D[1]; dz = C[2] - D[2];
   const double cd2 = dx * dx + dy * dy + dz * dz;
   const double ecd = -(zc * zd) / (zc + zd) * cd2;
   const double rsz = 1. / sqrt(z);
   g.Kfac = rsz * pref * gint_exp(eab + ecd);
   Cn[0] = Co[0] + kshift[0];
   Cn[1] = Co[1] + kshift[1];
   Cn[2] = Co[2] + kshift[2];
   Dn[0] = Co[0] + dimg[0] + neighs[n2 * 3 + 0] + kshift[0];
   Dn[1] = Co[1] + dimg[1] + neighs[n2 * 3 + 1] + kshift[1];
   Dn[2] = Co[2] + dimg[2] + neighs[n2 * 3 + 2] + kshift[2];
   Q[0] = (zc * Cn[0] + zd * Dn[0]) * inv_zcd;
   Q[1] = (zc * Cn[1] + zd * Dn[1]) * inv_zcd;
   Q[2] = (zc * Cn[2] + zd * Dn[2]) * inv_zcd;
   for (int k = 0; k < 3; k++) {
      g.P[k] = P[k];
      g.Q[k] = Q[k];
      g.PA[k] = P[k] - A[k];
      g.QC[k] = Q[k] - Cn[k];
   }
   g.zab = zab;
   g.zcd = zcd;
   g.rho = zab * zcd * inv_z;
   g.Rr = t.R_cut * sqrt(g.rho);
   g.i2ab = 0.5 * inv_zab;
   g.i2cd = 0.5 * inv_zcd;
}

// order-independent survivor checksum, same function as the oracle's
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long slot_hash(uint32_t set_ij, uint32_t set_kl, int n1, int n2, uint32_t ip,
                                                        int n3)
{
   const unsigned long long w0 = ((unsigned long long)(set_ij & 0xffff) << 32) | (set_ij >> 16);
   const unsigned long long w1 = ((unsigned long long)(set_kl & 0xffff) << 32) | (set_kl >> 16);
   const unsigned long long w2 = ((unsigned long long)n1 << 24) | ((unsigned long long)n2 << 16);
   const unsigned long long pk = ((ip & 15) << 12) | (((ip >> 4) & 15) << 8) | (((ip >> 8) & 15) << 4) | ((ip >> 12) & 15);
   const unsigned long long w3 = (pk << 8) | (unsigned long long)n3;
   unsigned long long z = mix64(w0 + 0x9E3779B97F4A7C15ULL);
   z = mix64(z ^ w1);
   z = mix64(z ^ w2);
   z = mix64(z ^ w3);
   return z;
}

// contraction coefficient product of combination l of class k
__device__ __forceinline__ double kprod(const TypePlan &tp, const GroupRec &g, const double *data, int k, int l,
                                        uint32_t ip)
{
   const int sh = tp.cls_sh[k];
   const int sa = sh & 15, sb = (sh >> 4) & 15, sc = (sh >> 8) & 15, sd = (sh >> 12) & 15;
   const int nlb = tp.nl[1][sb], nlc = tp.nl[2][sc], nld = tp.nl[3][sd];
   const int ia = l / (nlb * nlc * nld), ib = (l / (nlc * nld)) % nlb, ic = (l / nld) % nlc, id = l % nld;
   const int npa = (g.np4 & 15) + 1, npb = ((g.np4 >> 4) & 15) + 1, npc = ((g.np4 >> 8) & 15) + 1,
             npd = ((g.np4 >> 12) & 15) + 1;
   return data[g.Kof[0][sa] + ia * npa + (ip & 15)] * data[g.Kof[1][sb] + ib * npb + ((ip >> 4) & 15)] *
          data[g.Kof[2][sc] + ic * npc + ((ip >> 8) & 15)] * data[g.Kof[3][sd] + id * npd + ((ip >> 12) & 15)];
}

// the same product from the coefficients staged in LDS (layout co_off)
__device__ __forceinline__ double kprod_lds(const TypePlan &tp, const double *co, int k, int l, uint32_t ip)
{
   const int sh = tp.cls_sh[k];
   const int sa = sh & 15, sb = (sh >> 4) & 15, sc = (sh >> 8) & 15, sd = (sh >> 12) & 15;
   const int nlb = tp.nl[1][sb], nlc = tp.nl[2][sc], nld = tp.nl[3][sd];
   const int ia = l / (nlb * nlc * nld), ib = (l / (nlc * nld)) % nlb, ic = (l / nld) % nlc, id = l % nld;
   return co[tp.co_off[0][sa] + ia * MAXNP + (ip & 15)] * co[tp.co_off[1][sb] + ib * MAXNP + ((ip >> 4) & 15)] *
          co[tp.co_off[2][sc] + ic * MAXNP + ((ip >> 8) & 15)] * co[tp.co_off[3][sd] + id * MAXNP + ((ip >> 12) & 15)];
}

// offset in the type's block set of element j of a class's block set
// (blocks e in [la,LA], f in [lc,LC], e outer, layout [ie][jf] in a block)
__device__ __forceinline__ int class_src(const TypePlan &tp, int la, int LA, int lc, int LC, int j)
{
   for (int e = la; e <= LA; e++) {
      const int ne = d_nco(e);
      for (int f = lc; f <= LC; f++) {
         const int sz = ne * d_nco(f);
         if (j < sz) return tp.cpoff[e][f] + j;
         j -= sz;
      }
   }
   return 0;
}

// sum_j T2_j sum_{l<14-j} T1_l c[jl]: the operations of gint_cheb_sum in the
// same order (so the value is bitwise the host's), with T2_j produced by
// the recurrence of gint_cheb14 inside the loop to keep few values live
__device__ __forceinline__ double d_cheb_sum(const double *__restrict__ c, const double *T1, double x2)
{
   GINT_NO_CONTRACT
   double tmp = 0.0, tm1 = 0.0, tm2 = 0.0;
   int jl = 0;
#pragma unroll 7
   for (int j = 0; j < 14; j++) {
      double t2;
      if (j == 0) t2 = 1.0;
      else if (j == 1) t2 = GINT_SQRT2 * x2;
      else if (j == 2) t2 = 2. * x2 * tm1 - GINT_SQRT2;
      else t2 = 2. * x2 * tm1 - tm2;
      double dot = 0.0;
#pragma unroll
      for (int l = 0; l < 14; l++)
         if (l < 14 - j) dot += T1[l] * c[jl + l];
      tmp += dot * t2;
      jl += 14 - j;
      tm2 = tm1;
      tm1 = t2;
   }
   return tmp;
}

// ---------------------------------------------------------------------------
// K1: G_m -> Hermite R -> cell sum -> (e0|f0), one wavefront per primitive
// quartet.  FIX: one wavefront per listed (primitive, cell) slot, no
// screening, blocks and per-L Fm[0] into the fix-up workspace.
// ---------------------------------------------------------------------------
struct FmLayout {
   int SB;      // survivors per Hermite batch
   int o_fmv;   // per lane: G_m, then D_m = (-2 rho)^m G_m
   int o_sv;    // per lane: X Y Z of P-Q-g, TG1 (or T), TG2
   int o_R;     // Hermite layers, 2 x SB x nh
   int o_int;   // byte offset of 64 lane infos (branch | patch << 2) and 64 survivor lanes
   int bytes;
};

constexpr int GEO = 24;   // doubles of the primitive geometry at the start of LDS
enum { G_P = 0, G_Q = 3, G_PA = 6, G_QC = 9, G_KFAC = 12, G_RHO = 13, G_RR = 14, G_I2AB = 15, G_I2CD = 16 };

static FmLayout fm_layout(const TypePlan &tp, bool fix, bool pair = false)
{
   FmLayout y;
   const int L = tp.L, nh = tp.nh;
   const int geo = pair ? 2 * GEO : GEO;   // one geometry record per primitive quartet of the wavefront
   y.o_fmv = geo;
   y.o_sv = geo + 64 * (L + 1);
   y.o_R = y.o_sv + 320;
   // the Hermite batch fills what is left of a 10 KB budget: small enough
   // that 4 wavefronts per SIMD fit in LDS
   int SB = fix ? 1 : (1280 - y.o_R) / (2 * nh);
   SB = std::max(1, std::min(16, SB));
   y.SB = SB;
   const int cell_phase = y.o_R + 2 * SB * nh;
   const int e_phase = geo + nh + 6 * 28 + tp.nbh * tp.nkc;
   const int dbl = std::max(cell_phase, e_phase);
   y.o_int = 8 * dbl;
   y.bytes = y.o_int + 4 * 128;
   return y;
}

// at most 128 VGPRs (4 wavefronts per SIMD, with the LDS budget above): the
// kernel waits on memory half of its time, so wavefronts in flight pay more
// than the few values it then keeps in scratch
//
// PAIR (at most 32 cells, at most 256 Hermite functions): two primitive
// quartets per wavefront.  Lanes 0-31 take the cells of the first, lanes
// 32-63 those of the second; their survivors share the G_m and Hermite
// batches and each quartet keeps its own cell sum and transform.
template <bool FIX, bool PAIR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_fm_r(const TypePlan *__restrict__ tpp, DevTables t, BatchDev b, FmLayout lay)
{
   static_assert(!(FIX && PAIR), "the fix-up takes one slot per wavefront");
   constexpr int NPR = PAIR ? 2 : 1;      // primitive quartets per wavefront
   constexpr int NACC = PAIR ? 4 : 8;     // Hermite functions per lane and quartet, / 64
   extern __shared__ double lds[];
   const TypePlan &tp = *tpp;
   const int lane = threadIdx.x;
   const int half = PAIR ? lane >> 5 : 0;   // the lane's quartet
   const int L = tp.L, nh = tp.nh, Lp1 = L + 1;
   const int SB = lay.SB;
   double *G = lds + GEO * half;      // geometry of the lane's quartet
   double *fmv = lds + lay.o_fmv;     // [64][L+1]
   double *sv = lds + lay.o_sv;       // [64][5]
   double *Rb = lds + lay.o_R;        // [2][SB][nh]
   int *svi = (int *)((char *)lds + lay.o_int);
   int *svl = svi + 64;
   const uint32_t nwork = FIX ? min(b.fix_count[0], b.fix_cap) : (b.p1 - b.p0);
   const uint32_t *__restrict__ e1 = b.e1, *__restrict__ e2 = b.e2;

   for (uint32_t w0 = blockIdx.x * NPR; w0 < nwork; w0 += gridDim.x * NPR) {
      const uint32_t w = w0 + half;
      const bool hvalid = w < nwork;                    // the lane's quartet exists
      const int nprim = (int)min((uint32_t)NPR, nwork - w0);
      uint32_t p;
      int fix_n3 = 0;
      if (FIX) {
         p = b.fix[2 * w];
         fix_n3 = (int)(b.fix[2 * w + 1] & 0xff);
      } else {
         p = b.p0 + (hvalid ? w : w0);
      }
      const uint32_t r = b.prm_row[p];
      const RowRec row = b.rows[r];
      const GroupRec &g = b.groups[row.group];
      const uint32_t ip = b.prm_ip[p];
      const int n1 = row.n12 & 0xffff, n2 = row.n12 >> 16;
      if ((lane & (64 / NPR - 1)) == 0 && hvalid) {
         const double *data = b.data;
         PrimGeo geo;
         prim_prepare(data + g.R[0], data + g.R[1], data + g.R[2], data + g.R[3], data[g.Z[0] + (ip & 15)],
                      data[g.Z[1] + ((ip >> 4) & 15)], data[g.Z[2] + ((ip >> 8) & 15)],
                      data[g.Z[3] + ((ip >> 12) & 15)], n1, n2, t, geo);
         for (int k = 0; k < 3; k++) {
            G[G_P + k] = geo.P[k];
            G[G_Q + k] = geo.Q[k];
            G[G_PA + k] = geo.PA[k];
            G[G_QC + k] = geo.QC[k];
         }
         G[G_KFAC] = geo.Kfac;
         G[G_RHO] = geo.rho;
         G[G_RR] = geo.Rr;
         G[G_I2AB] = geo.i2ab;
         G[G_I2CD] = geo.i2cd;
      }
      __syncthreads();
      double acc[NPR][NACC];
#pragma unroll
      for (int q = 0; q < NPR; q++)
#pragma unroll
         for (int i = 0; i < NACC; i++) acc[q][i] = 0.0;
      double quirk = 0.0;
      unsigned long long n_surv = 0, csum = 0, n_any = 0;
      bool any_surv[NPR];
#pragma unroll
      for (int q = 0; q < NPR; q++) any_surv[q] = false;
      const unsigned all = (1u << tp.nL) - 1u;
      const int ncell = FIX ? 1 : t.Ng;

      for (int base = 0; base < ncell; base += 64) {
         const int n3 = FIX ? fix_n3 : base + (PAIR ? (lane & 31) : lane);
         const bool valid = FIX ? (lane == 0) : (hvalid && n3 < t.Ng);
         int status = 0;  // 1 kept by every class, 2 screened by every class, 3 class-dependent
         unsigned mask = 0;
         double *myf = fmv + lane * Lp1;
         if (valid) {
            const double Kfac = G[G_KFAC], rho = G[G_RHO];
            double X[3], T, Fm0 = 0.0, tg1 = 0.0, tg2 = 0.0;
            int br, patch = 0;
            {
               GINT_NO_CONTRACT
               X[0] = G[G_P + 0] - G[G_Q + 0] - t.neighs[n3 * 3 + 0];
               X[1] = G[G_P + 1] - G[G_Q + 1] - t.neighs[n3 * 3 + 1];
               X[2] = G[G_P + 2] - G[G_Q + 2] - t.neighs[n3 * 3 + 2];
               const double rpq2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
               T = rho * rpq2;
            }
            br = t.truncated ? gint_tcg_locate(G[G_RR], T, t.grid_lo, t.grid_hi, t.bw, &patch, &tg1, &tg2) : GINT_G_BOYS;
            if (br == GINT_G_PATCH) {
               double T1[14];
               gint_cheb14(tg1, T1);
               {
                  GINT_NO_CONTRACT
                  Fm0 = d_cheb_sum(t.C0 + (size_t)t.ld_C0 * patch, T1, tg2) * Kfac;
               }
               myf[0] = Fm0;
            } else if (br == GINT_G_BOYS) {
               gint_boys(L, T, myf, t.ftable, Kfac);
               Fm0 = myf[0];
            }
            const bool lsens = br == GINT_G_BOYS && T >= 1.e-13 && T < 12.;   // F_0 depends on the start order
            if (br == GINT_G_ZERO) {
               status = 2;
            } else {
               mask = Fm0 > 1.e-12 ? all : 0u;
               // F_0 of the downward recursion depends on the starting order
               // by rounding only; decide each class L exactly this close to
               // the threshold
               if (lsens && tp.nL > 1 && fabs(Fm0 - 1.e-12) <= 1.e-20) {
                  mask = 0u;
                  for (int i = 0; i < tp.nL; i++) {
                     gint_boys(tp.Ls[i], T, myf, t.ftable, Kfac);
                     if (myf[0] > 1.e-12) mask |= 1u << i;
                  }
                  gint_boys(L, T, myf, t.ftable, Kfac);
               }
               status = mask == all ? 1 : (mask == 0u ? 2 : 3);
            }
            if (FIX) {
               // per-L Fm[0] of the slot, for the classes that screen it
               double *f0 = b.fix_ws + (size_t)w * (tp.ncp + 16) + tp.ncp;
               for (int i = 0; i < tp.nL; i++) {
                  double v = Fm0;
                  if (lsens) {
                     gint_boys(tp.Ls[i], T, myf, t.ftable, Kfac);
                     v = myf[0];
                  }
                  f0[i] = v;
               }
               if (lsens) gint_boys(L, T, myf, t.ftable, Kfac);
               status = br == GINT_G_ZERO ? 2 : 1;
            } else if (b.force_fix && status == 1) {
               status = 3;
            }
            if (!FIX) {
               if (status == 2) {
                  if (tp.has00) quirk += Fm0;
               } else {
                  if (b.stats) {
                     const unsigned long long h = slot_hash(g.set_ij, g.set_kl, n1, n2, ip, n3);
                     if (mask == all) {
                        n_surv += (unsigned long long)tp.nclass;
                        csum += h * tp.mult_all;
                     } else {
                        unsigned long long cm = 0;
                        for (int k = 0; k < tp.nclass; k++)
                           if ((mask >> tp.cls_gate[k]) & 1u) {
                              n_surv++;
                              cm += tp.cls_mult[k];
                           }
                        csum += h * cm;
                     }
                  }
                  if (status == 3) {
                     const uint32_t idx = atomicAdd(&b.fix_count[0], 1u);
                     if (idx < b.fix_cap) {
                        b.fix[2 * idx] = p;
                        b.fix[2 * idx + 1] = (uint32_t)n3 | (mask << 8);
                     } else {
                        // list full: every class keeps the slot (its G_0 is
                        // within rounding of the 1e-12 threshold)
                        atomicAdd(&b.fix_count[1], 1u);
                        status = 1;
                        n_any++;
                     }
                  } else {
                     n_any++;
                  }
               }
            }
            if (status == 1) {
               sv[lane * 5 + 0] = X[0];
               sv[lane * 5 + 1] = X[1];
               sv[lane * 5 + 2] = X[2];
               sv[lane * 5 + 3] = br == GINT_G_PATCH ? tg1 : T;
               sv[lane * 5 + 4] = tg2;
               svi[lane] = br | (patch << 2);
            }
         }
         // survivors of this pass, in lane order
         const unsigned long long smask = __ballot(status == 1);
         const int nsv = __popcll(smask);
         const int nsv0 = PAIR ? __popcll(smask & 0xffffffffull) : nsv;   // survivors of the first quartet
         if (status == 1) svl[__popcll(smask & ((1ull << lane) - 1ull))] = lane;
         __syncthreads();
         if (nsv == 0) continue;
         any_surv[0] = any_surv[0] || nsv0 > 0;
         if (PAIR) any_surv[NPR - 1] = any_surv[NPR - 1] || nsv > nsv0;
         // G_1..G_L of the patch survivors, one (survivor, order) per lane
         {
            for (int idx = lane; idx < nsv * L; idx += 64) {
               const int s = idx / L, m = idx - s * L + 1;
               const int ln = svl[s];
               const double Kfac = lds[GEO * (PAIR ? ln >> 5 : 0) + G_KFAC];
               const int info = svi[ln];
               if ((info & 3) == GINT_G_PATCH) {
                  double T1[14];
                  gint_cheb14(sv[ln * 5 + 3], T1);
                  const double *c = t.C0 + (size_t)t.ld_C0 * (info >> 2) + m * TCG_NCOEF;
                  fmv[ln * Lp1 + m] = d_cheb_sum(c, T1, sv[ln * 5 + 4]) * Kfac;
               }
            }
         }
         __syncthreads();
         {
            for (int idx = lane; idx < nsv * Lp1; idx += 64) {
               const int s = idx / Lp1, m = idx - s * Lp1;
               const int ln = svl[s];
               const double m2rho = -2.0 * lds[GEO * (PAIR ? ln >> 5 : 0) + G_RHO];
               double pw = 1.0;
               for (int k = 0; k < m; k++) pw *= m2rho;
               fmv[ln * Lp1 + m] *= pw;
            }
         }
         __syncthreads();
         // Hermite recursion R^(n)_tuv = coef R^(n+1)_{tuv-2d} + X_d R^(n+1)_{tuv-d}
         for (int s0 = 0; s0 < nsv; s0 += SB) {
            const int nb = min(SB, nsv - s0);
            double *prv = Rb, *cur = Rb + SB * nh;
            if (lane < nb) prv[lane * nh] = fmv[svl[s0 + lane] * Lp1 + L];
            __syncthreads();
            for (int n = L - 1; n >= 1; n--) {
               const int mn = d_nherm(L - n);
               for (int h = lane; h < mn; h += 64) {
                  if (h == 0) {
                     for (int s = 0; s < nb; s++) cur[s * nh] = fmv[svl[s0 + s] * Lp1 + n];
                  } else {
                     const HermEntry e = c_herm.e[h];
                     const double cf = (double)e.coef;
                     for (int s = 0; s < nb; s++) {
                        double v = sv[svl[s0 + s] * 5 + e.dir] * prv[s * nh + e.m1];
                        if (e.m2 >= 0) v += cf * prv[s * nh + e.m2];
                        cur[s * nh + h] = v;
                     }
                  }
               }
               __syncthreads();
               double *tmp = prv;
               prv = cur;
               cur = tmp;
            }
            // the n = 0 layer goes straight into the cell sum of its quartet
#pragma unroll
            for (int i = 0; i < NACC; i++) {
               const int h = lane + 64 * i;
               if (h < nh) {
                  double a0 = acc[0][i], a1 = acc[NPR - 1][i];
                  if (h == 0) {
                     for (int s = 0; s < nb; s++) {
                        const double v = fmv[svl[s0 + s] * Lp1];
                        if (!PAIR || s0 + s < nsv0) a0 += v;
                        else a1 += v;
                     }
                  } else {
                     const HermEntry e = c_herm.e[h];
                     const double cf = (double)e.coef;
                     for (int s = 0; s < nb; s++) {
                        double v = sv[svl[s0 + s] * 5 + e.dir] * prv[s * nh + e.m1];
                        if (e.m2 >= 0) v += cf * prv[s * nh + e.m2];
                        if (!PAIR || s0 + s < nsv0) a0 += v;
                        else a1 += v;
                     }
                  }
                  acc[0][i] = a0;
                  if (PAIR) acc[NPR - 1][i] = a1;
               }
            }
            __syncthreads();
         }
      }

      // ---- Hermite -> cartesian, once per primitive quartet ----
      double *sR = lds + lay.o_fmv;      // [nh]
      double *Ex = sR + nh;              // [3][28] bra E^{e0}_t
      double *Fk = Ex + 84;              // [3][28] ket (-1)^tau E^{f0}_tau
      double *W = Fk + 84;               // [nbh][nkc]
      // screened slots' Fm[0], summed per quartet (lane 0, and lane 32 for the second)
      for (int off = 32 / NPR; off > 0; off >>= 1) quirk += __shfl_down(quirk, off, 64 / NPR);
      double qsum[NPR];
#pragma unroll
      for (int q = 0; q < NPR; q++) qsum[q] = __shfl(quirk, 32 * q, 64);
#pragma unroll
      for (int q = 0; q < NPR; q++) {
         if (q >= nprim) break;
         const double *Gq = lds + GEO * q;
         double *dst = FIX ? b.fix_ws + (size_t)w * (tp.ncp + 16) : b.ac0 + (size_t)(w0 + q) * tp.ncp;
         // calibration of the mixed-precision routing: the largest block
         // element of each set quartet (a positive double orders as its bits)
         unsigned long long *gm = nullptr;
         if (!FIX && b.gmag) gm = (unsigned long long *)(b.gmag + b.rows[b.prm_row[b.p0 + w0 + q]].group);
         if (!FIX && !any_surv[q]) {
            // no cell survives: the blocks are zero but for the screened
            // slots' Fm[0] in (00|00)
            for (int j = lane; j < tp.ncp; j += 64) dst[j] = (j == 0 && tp.has00) ? qsum[q] : 0.0;
            if (gm && lane == 0 && tp.has00) atomicMax(gm, (unsigned long long)__double_as_longlong(fabs(qsum[q])));
            __syncthreads();
            continue;
         }
         __syncthreads();
#pragma unroll
         for (int i = 0; i < NACC; i++) {
            const int h = lane + 64 * i;
            if (h < nh) sR[h] = acc[q][i];
         }
         if (lane < 6) {
            const int d = lane % 3;
            const bool ket = lane >= 3;
            const int Lm = ket ? tp.LC : tp.LA;
            const double x = ket ? Gq[G_QC + d] : Gq[G_PA + d];
            const double i2 = ket ? Gq[G_I2CD] : Gq[G_I2AB];
            double *E = (ket ? Fk : Ex) + d * 28;
            E[0] = 1.0;
            for (int e = 0; e < Lm; e++) {
               const double *Ee = E + e * (e + 1) / 2;
               double *En = E + (e + 1) * (e + 2) / 2;
               for (int tt = 0; tt <= e + 1; tt++) {
                  double v = 0.0;
                  if (tt > 0) v += i2 * Ee[tt - 1];
                  if (tt <= e) v += x * Ee[tt];
                  if (tt + 1 <= e) v += (double)(tt + 1) * Ee[tt + 1];
                  En[tt] = v;
               }
            }
            if (ket)
               for (int e = 0; e <= Lm; e++)
                  for (int tt = 1; tt <= e; tt += 2) E[e * (e + 1) / 2 + tt] = -E[e * (e + 1) / 2 + tt];
         }
         __syncthreads();
         const int nkc = tp.nkc;
         // W[tuv][f] = sum E^f_tau E^f_nu E^f_phi R[t+tau, u+nu, v+phi]
         for (int idx = lane; idx < tp.nbh * nkc; idx += 64) {
            const uint32_t it = e1[idx];
            const int tt = it & 7, uu = (it >> 3) & 7, vv = (it >> 6) & 7;
            const int fx = (it >> 9) & 7, fy = (it >> 12) & 7, fz = (it >> 15) & 7;
            const double *Fx = Fk + fx * (fx + 1) / 2, *Fy = Fk + 28 + fy * (fy + 1) / 2, *Fz = Fk + 56 + fz * (fz + 1) / 2;
            double s = 0.0;
            for (int a = 0; a <= fx; a++)
               for (int c = 0; c <= fy; c++) {
                  const double cxy = Fx[a] * Fy[c];
                  const int t1 = tt + a, u1 = uu + c;
                  for (int d = 0; d <= fz; d++) {
                     const int v1 = vv + d, N = t1 + u1 + v1;
                     const int H = N * (N + 1) * (N + 2) / 6 + (N - t1) * (N - t1 + 1) / 2 + v1;
                     s += cxy * Fz[d] * sR[H];
                  }
               }
            W[it >> 18] = s;
         }
         __syncthreads();
         // (e0|f0) = sum E^e_t E^e_u E^e_v W[tuv][f]
         double vmax = 0.0;
         for (int j = lane; j < tp.ncp; j += 64) {
            const uint32_t it = e2[j];
            const int ex = it & 7, ey = (it >> 3) & 7, ez = (it >> 6) & 7, col = (int)((it >> 9) & 127);
            const double *Ea = Ex + ex * (ex + 1) / 2, *Eb = Ex + 28 + ey * (ey + 1) / 2, *Ec = Ex + 56 + ez * (ez + 1) / 2;
            double s = 0.0;
            for (int a = 0; a <= ex; a++)
               for (int c = 0; c <= ey; c++) {
                  const double cxy = Ea[a] * Eb[c];
                  for (int d = 0; d <= ez; d++) {
                     const int N = a + c + d;
                     const int H = N * (N + 1) * (N + 2) / 6 + (N - a) * (N - a + 1) / 2 + d;
                     s += cxy * Ec[d] * W[H * nkc + col];
                  }
               }
            const int jd = (int)(it >> 16);
            if (!FIX && jd == 0 && tp.has00) s += qsum[q];
            dst[jd] = s;
            vmax = fmax(vmax, fabs(s));
         }
         if (gm) {
            for (int off = 32; off > 0; off >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
            if (lane == 0) atomicMax(gm, (unsigned long long)__double_as_longlong(vmax));
         }
         __syncthreads();
      }

      if (!FIX && b.stats) {
         for (int off = 32; off > 0; off >>= 1) {
            n_surv += __shfl_down(n_surv, off, 64);
            csum += __shfl_down(csum, off, 64);
            n_any += __shfl_down(n_any, off, 64);
         }
         if (lane == 0) {
            // per-workgroup partials spread over STATS_SLOTS counters: one
            // address hit by every workgroup of the chip serialises them
            unsigned long long *st = b.stats + 3 * (blockIdx.x & (STATS_SLOTS - 1));
            if (n_surv) {
               atomicAdd(&st[0], n_surv);
               atomicAdd(&st[1], csum);
            }
            // algorithmic FLOPs: the reference chain's Fm and VRR of every
            // (slot, class) pair kept (reference src/compute_Fm.cpp,
            // src/compute_VRR.cpp; summed over the type's classes)
            atomicAdd(&st[2], n_any * (unsigned long long)tp.ref_slot_flops);
         }
      }
   }
}

// ---------------------------------------------------------------------------
// K2: per (set quartet, class), for a batch of its contractions at once:
// contraction over each row's primitive quartets (reference
// src/compute_ECO.cpp:12-214), HRR of the row (src/compute_HRR.cpp), sum
// over the rows, cartesian -> spherical (src/compute_SPH.cpp:58-237) and
// the exchange update of K (src/compute_KS.cpp:8-117), reduced over the set
// quartet's contractions before the atomics when the spherical integrals of
// all contractions fit on chip.
// ---------------------------------------------------------------------------
constexpr int CPL_BYTES = (sizeof(ClassPlan) + 15) & ~15;   // LDS copy of the class plan
constexpr int CQ_PCH = 8;    // primitive quartets per contraction step
// LDS copies of the tables that the HRR and the spherical transform index per
// lane (from constant memory they would be vector loads): c2s (110 doubles),
// then the cartesian first direction [4][10], minus [4][10][3] and plus
// [7][28][3] tables
constexpr int CQ_TAB_C2S = 0, CQ_TAB_FIRST = 8 * GINT_C2S_SIZE, CQ_TAB_MINUS = CQ_TAB_FIRST + 40,
              CQ_TAB_PLUS = CQ_TAB_MINUS + 120, CQ_TAB_BYTES = (CQ_TAB_PLUS + 588 + 15) & ~15;
// the classes of a k_gqs part, listed in LDS after the tables
constexpr int CQ_PLIST_BYTES = (4 * (MAXTC + 1) + 15) & ~15;


__device__ __forceinline__ double *lds_base()
{
   extern __shared__ double lds[];
   return lds;
}

// x / d for 0 <= x < 2^22, d >= 1: float reciprocal estimate, one correction
__device__ __forceinline__ int fdiv(int x, int d)
{
   int q = (int)((float)x * __builtin_amdgcn_rcpf((float)d));
   const int rem = x - q * d;
   if (rem < 0) q--;
   else if (rem >= d) q++;
   return q;
}

__device__ void row_ab_cd(const GroupRec &g, const RowRec &row, const double *data, const DevTables &t, double AB[3],
                          double CD[3])
{
   const double *Ao = data + g.R[0], *Bo = data + g.R[1], *Co = data + g.R[2], *Do = data + g.R[3];
   const int n1 = row.n12 & 0xffff, n2 = row.n12 >> 16;
   double ABs[3], CDs[3];
   gint_pbc(Ao, Bo, t.cell, ABs);
   gint_pbc(Co, Do, t.cell, CDs);
   for (int k = 0; k < 3; k++) {
      const double B = Ao[k] + ABs[k] + t.neighs[n1 * 3 + k];
      const double D = Co[k] + CDs[k] + t.neighs[n2 * 3 + k];
      AB[k] = Ao[k] - B;
      CD[k] = Co[k] - D;
   }
}

// block offsets of contraction l in an affine group
__device__ __forceinline__ void ks_offsets(const KsGroup &ks, int nlb, int nlc, int nld, int l, size_t off[4])
{
   const int ia = l / (nlb * nlc * nld), ib = (l / (nlc * nld)) % nlb, ic = (l / nld) % nlc, id = l % nld;
   off[0] = (size_t)((long long)ks.base[0] + (long long)ia * ks.step[0] + (long long)ic * ks.step[1]);
   off[1] = (size_t)((long long)ks.base[1] + (long long)ia * ks.step[2] + (long long)id * ks.step[3]);
   off[2] = (size_t)((long long)ks.base[2] + (long long)ib * ks.step[4] + (long long)ic * ks.step[5]);
   off[3] = (size_t)((long long)ks.base[3] + (long long)ib * ks.step[6] + (long long)id * ks.step[7]);
}

__device__ __forceinline__ unsigned blk_idx(bool T, unsigned ld, unsigned row_i, unsigned col_i)
{
   return T ? row_i * ld + col_i : col_i * ld + row_i;
}

// symmetry factor of a K layout code: 2, 1, 1/2, 1/4, or 0 (code 4)
__device__ __forceinline__ double symm_code_fac(uint32_t c) { return c >= 4 ? 0.0 : (c == 0 ? 2.0 : (c == 1 ? 1.0 : (c == 2 ? 0.5 : 0.25))); }

// K accumulators: the K updates of a set quartet's contractions summed in
// LDS instead of four atomics per integral.  Per spin the accumulators are
// the four K blocks, [ac: NA x NC][ad: NA x ND][bc: NB x NC][bd: NB x ND],
// rows and columns (contraction, spherical function) of their centres.  Each
// element is the sum, over the two other spherical indices, of the
// integrals times the P block it pairs with (the reference's updates
// K_bd += I P_ac, K_bc += I P_ad, K_ad += I P_bc, K_ac += I P_bd per
// integral, src/compute_KS.cpp:8-117, regrouped; the factor -f is applied
// at the flush).
//
// kacc_add for the nb contractions l0 .. l0 + nb - 1 of a pass at once:
// the (contraction, element) items spread over all threads, each sum added
// to its accumulator with an LDS atomic (contractions of a pass that share
// two of the indices meet on the same element).  I(l, isa, isb, isc, isd):
// the integrals of contraction l0 + l.
template <int NT, typename IntegralFn>
__device__ __forceinline__ void kacc_add_pass(const KTargets &kt, double *acc, int kacc, const KsGroup &ks, int l0,
                                              int nb, int nlb, int nlc, int nld, int nsa, int nsb, int nsc, int nsd,
                                              int NA, int NB, int NC, int ND, IntegralFn I)
{
   const unsigned ldac = ks.ld01 & 0xffff, ldad = ks.ld01 >> 16, ldbc = ks.ld23 & 0xffff, ldbd = ks.ld23 >> 16;
   const bool Tac = ks.flags & 1, Tad = ks.flags & 2, Tbc = ks.flags & 4, Tbd = ks.flags & 8;
   const int n_bd = nsb * nsd, n_bc = nsb * nsc, n_ad = nsa * nsd, n_ac = nsa * nsc;
   const int ntask = n_bd + n_bc + n_ad + n_ac;
   for (int it = threadIdx.x; it < nb * ntask; it += NT) {
      const int l = it / ntask, t = it - l * ntask, q = l0 + l;
      const int ia = q / (nlb * nlc * nld), ib = (q / (nlc * nld)) % nlb, ic = (q / nld) % nlc, id = q % nld;
      size_t off[4];
      ks_offsets(ks, nlb, nlc, nld, q, off);
      for (int sp = 0; sp < kt.nspin; sp++) {
         const double *P = kt.P[sp];
         double *A = acc + sp * kacc;
         double sum = 0.0;
         int dst;
         if (t < n_bd) {
            const int isb = t / nsd, isd = t - isb * nsd;
            for (int isa = 0; isa < nsa; isa++)
               for (int isc = 0; isc < nsc; isc++)
                  sum += I(l, isa, isb, isc, isd) * P[off[0] + blk_idx(Tac, ldac, isa, isc)];
            dst = NA * NC + NA * ND + NB * NC + (ib * nsb + isb) * ND + id * nsd + isd;
         } else if (t < n_bd + n_bc) {
            const int u = t - n_bd, isb = u / nsc, isc = u - isb * nsc;
            for (int isa = 0; isa < nsa; isa++)
               for (int isd = 0; isd < nsd; isd++)
                  sum += I(l, isa, isb, isc, isd) * P[off[1] + blk_idx(Tad, ldad, isa, isd)];
            dst = NA * NC + NA * ND + (ib * nsb + isb) * NC + ic * nsc + isc;
         } else if (t < n_bd + n_bc + n_ad) {
            const int u = t - n_bd - n_bc, isa = u / nsd, isd = u - isa * nsd;
            for (int isb = 0; isb < nsb; isb++)
               for (int isc = 0; isc < nsc; isc++)
                  sum += I(l, isa, isb, isc, isd) * P[off[2] + blk_idx(Tbc, ldbc, isb, isc)];
            dst = NA * NC + (ia * nsa + isa) * ND + id * nsd + isd;
         } else {
            const int u = t - n_bd - n_bc - n_ad, isa = u / nsc, isc = u - isa * nsc;
            for (int isb = 0; isb < nsb; isb++)
               for (int isd = 0; isd < nsd; isd++)
                  sum += I(l, isa, isb, isc, isd) * P[off[3] + blk_idx(Tbd, ldbd, isb, isd)];
            dst = (ia * nsa + isa) * NC + ic * nsc + isc;
         }
         atomicAdd(&A[dst], sum);
      }
   }
}

// the K accumulators into K: one atomic per element of the four blocks
template <int NT>
__device__ void kacc_flush(const KTargets &kt, const KsGroup &ks, const double *acc, int kacc, int nsa, int nsb, int nsc,
                           int nsd, int NA, int NB, int NC, int ND)
{
   const double fac = symm_code_fac((ks.flags >> 4) & 7) * kt.hf_fac;
   if (fac == 0.0) return;
   const unsigned ldac = ks.ld01 & 0xffff, ldad = ks.ld01 >> 16, ldbc = ks.ld23 & 0xffff, ldbd = ks.ld23 >> 16;
   const bool Tac = ks.flags & 1, Tad = ks.flags & 2, Tbc = ks.flags & 4, Tbd = ks.flags & 8;
   for (int sp = 0; sp < kt.nspin; sp++) {
      double *K = kt.K[sp];
      for (int e = threadIdx.x; e < kacc; e += NT) {
         const double v = acc[sp * kacc + e];
         if (v == 0.0) continue;
         long long o;
         if (e < NA * NC) {
            const int r = e / NC, c = e - r * NC, i = r / nsa, si = r - i * nsa, j = c / nsc, sj = c - j * nsc;
            o = (long long)ks.base[0] + (long long)i * ks.step[0] + (long long)j * ks.step[1] + blk_idx(Tac, ldac, si, sj);
         } else if (e < NA * NC + NA * ND) {
            const int u = e - NA * NC, r = u / ND, c = u - r * ND, i = r / nsa, si = r - i * nsa, j = c / nsd,
                      sj = c - j * nsd;
            o = (long long)ks.base[1] + (long long)i * ks.step[2] + (long long)j * ks.step[3] + blk_idx(Tad, ldad, si, sj);
         } else if (e < NA * NC + NA * ND + NB * NC) {
            const int u = e - NA * NC - NA * ND, r = u / NC, c = u - r * NC, i = r / nsb, si = r - i * nsb, j = c / nsc,
                      sj = c - j * nsc;
            o = (long long)ks.base[2] + (long long)i * ks.step[4] + (long long)j * ks.step[5] + blk_idx(Tbc, ldbc, si, sj);
         } else {
            const int u = e - NA * NC - NA * ND - NB * NC, r = u / ND, c = u - r * ND, i = r / nsb, si = r - i * nsb,
                      j = c / nsd, sj = c - j * nsd;
            o = (long long)ks.base[3] + (long long)i * ks.step[6] + (long long)j * ks.step[7] + blk_idx(Tbd, ldbd, si, sj);
         }
         unsafeAtomicAdd(&K[(size_t)o], -fac * v);
      }
   }
}

// one (set quartet, class): the work of k_gqs for class cls_k of group gi
template <int NT>
__device__ void cq_one(const ClassPlan &cp, int cls_k, const TypePlan &tp, const DevTables &t, const BatchDev &b,
                       const KTargets &kt, const CqLayout &y, double *WS, const int *__restrict__ srcmap, uint32_t gi,
                       const double *co, uint32_t *ipb, const char *tabs)
{
   const unsigned char *t_first = (const unsigned char *)(tabs + CQ_TAB_FIRST);
   const signed char *t_minus = (const signed char *)(tabs + CQ_TAB_MINUS);
   const signed char *t_plus = (const signed char *)(tabs + CQ_TAB_PLUS);
   extern __shared__ double lds[];
   const int tid = threadIdx.x;
   const int cap = cp.ycap, lws = y.lws;
   double *Kp = lds + y.o_kp;
   double *Iall = lds + y.o_I;
   const int la = cp.la, lb = cp.lb, lc = cp.lc, ld = cp.ld, LC = cp.LC;
   const int Nc = cp.Nc, Ns = cp.Ns, ncp = cp.ncp, ncpt = tp.ncp;
   const int ncd = d_nco(lc) * d_nco(ld);
   const int sh = tp.cls_sh[cls_k];
   const int nla = tp.nl[0][sh & 15], nlb = tp.nl[1][(sh >> 4) & 15], nlc = tp.nl[2][(sh >> 8) & 15],
             nld = tp.nl[3][(sh >> 12) & 15];
   const int nq_all = nla * nlb * nlc * nld;
   const int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1, nsd = 2 * ld + 1;
   const int nscd = nsc * nsd, nsbcd = nsb * nscd;
   const double symm[5] = {2.0, 1.0, 0.5, 0.25, 0.0};
   const uint32_t nfix = min(b.fix_count[0], b.fix_cap);
   const int gate = tp.cls_gate[cls_k];
   const GroupRec &g = b.groups[gi];
   const KsGroup ks = group_ks(b, tp, g, cls_k);
   const int nq = (int)(ks.flags >> 16);
   if (nq == 0) return;
   const bool expl = (ks.flags & KS_EXPLICIT) != 0;
   const bool out_mode = (ks.flags & KS_OUTPUT) != 0;
   const bool reduce = y.reduce && !expl && !out_mode && nq == nq_all;
   // K accumulators in LDS when the integrals of all contractions do not fit
   const bool use_acc = !reduce && y.o_kacc >= 0 && !expl && !out_mode && nq == nq_all && kt.nspin * y.kacc <= y.kacc_cap;
   const int NA = nla * nsa, NB = nlb * nsb, NC = nlc * nsc, ND = nld * nsd;
   double *acc = lds + y.o_kacc;
   if (use_acc)
      for (int i = tid; i < kt.nspin * y.kacc; i += NT) acc[i] = 0.0;
   for (int l0 = 0; l0 < nq; l0 += y.lb) {
      const int nb = min(y.lb, nq - l0);
      __syncthreads();
      for (int idx = tid; idx < nb * Nc; idx += NT) {
         const int l = fdiv(idx, Nc);
         WS[l * lws + 3 * cap + (idx - l * Nc)] = 0.0;
      }
      for (uint32_t r = g.row0; r < g.row1; r++) {
         const RowRec row = b.rows[r];
         // ---- contraction of the row's primitive quartets into buf0 ----
         if (y.eco) {
            // done on the matrix cores (k_eco): the row's contracted blocks
            const double *E = b.eco + (size_t)(r - b.r0) * b.eco_row + y.eco_off + (size_t)l0 * ncp;
            for (int idx = tid; idx < nb * ncp; idx += NT) {
               const int l = fdiv(idx, ncp);
               WS[l * lws + (idx - l * ncp)] = E[idx];
            }
         }
         const uint32_t *IP = b.prm_ip + row.prm_off;
         const double *A0 = b.ac0 + (size_t)(row.prm_off - b.p0) * ncpt;
         if (!y.eco)
            for (int idx = tid; idx < nb * ncp; idx += NT) {
               const int l = fdiv(idx, ncp);
               WS[l * lws + (idx - l * ncp)] = 0.0;
            }
         for (int pc = 0; pc < (y.eco ? 0 : (int)row.n_prm); pc += CQ_PCH) {
            const int np = min(CQ_PCH, (int)row.n_prm - pc);
            __syncthreads();
            if (tid < np) ipb[tid] = IP[pc + tid];
            __syncthreads();
            for (int idx = tid; idx < nb * CQ_PCH; idx += NT) {
               const int l = idx / CQ_PCH, pp = idx % CQ_PCH;
               Kp[idx] = pp < np ? kprod_lds(tp, co, cls_k, l0 + l, ipb[pp]) : 0.0;
            }
            __syncthreads();
            const double *A = A0 + (size_t)pc * ncpt;
            // thread (g, j): column j of the step's blocks, loaded once, for the
            // contractions g, g + nlg, g + 2 nlg, ...
            const int nlg = max(1, min(NT / ncp, nb));
            const int per = (nb + nlg - 1) / nlg;
            for (int t = tid; t < nlg * ncp; t += NT) {
               const int g = fdiv(t, ncp), j = t - g * ncp;
               const int src = srcmap[j];
               double a[CQ_PCH];
#pragma unroll
               for (int pp = 0; pp < CQ_PCH; pp++) a[pp] = pp < np ? A[(size_t)pp * ncpt + src] : 0.0;
               for (int k = 0; k < per; k++) {
                  const int l = g + k * nlg;
                  if (l >= nb) break;
                  const double *kp = Kp + l * CQ_PCH;
                  double s = 0.0;
#pragma unroll
                  for (int pp = 0; pp < CQ_PCH; pp++) s += kp[pp] * a[pp];
                  WS[l * lws + j] += s;
               }
            }
         }
         // slots of this row whose screening depends on the class
         if (nfix) {
            __syncthreads();
            for (uint32_t e = 0; e < nfix; e++) {
               const uint32_t p = b.fix[2 * e];
               if (b.prm_row[p] != r) continue;
               const bool keep = ((b.fix[2 * e + 1] >> 8) >> gate) & 1u;
               const double *src = b.fix_ws + (size_t)e * (ncpt + 16);
               const uint32_t ip = b.prm_ip[p];
               if (keep) {
                  for (int idx = tid; idx < nb * ncp; idx += NT) {
                     const int l = fdiv(idx, ncp), j = idx - l * ncp;
                     WS[l * lws + j] += kprod(tp, g, b.data, cls_k, l0 + l, ip) * src[srcmap[j]];
                  }
               } else if (la == 0 && lc == 0) {
                  // a screened slot keeps its Fm[0] (reference src/compute_VRR.cpp:384-385)
                  for (int l = tid; l < nb; l += NT)
                     WS[l * lws] += kprod(tp, g, b.data, cls_k, l0 + l, ip) * src[ncpt + gate];
               }
               __syncthreads();
            }
         }
         __syncthreads();
         double AB[3], CD[3];
         row_ab_cd(g, row, b.data, t, AB, CD);
         int srcb = 0;
         // ket HRR (d -> c): layer dd from layer dd-1, blocks (e,c) e outer
         for (int dd = 1; dd <= ld; dd++) {
            const int dst = (srcb + 1) % 3;
            const int nd = d_nco(dd), ndm = d_nco(dd - 1);
            const int sz = cp.ksize[dd];
            for (int idx1 = tid; idx1 < nb * sz; idx1 += NT) {
               const int l = fdiv(idx1, sz), idx0 = idx1 - l * sz;
               const double *in = WS + l * lws + srcb * cap;
               int e = la, c = lc;
               for (;;) {
                  const int end_ec = cp.koff[dd][e][c] + d_nco(e) * d_nco(c) * nd;
                  if (idx0 < end_ec) break;
                  if (++c > LC - dd) {
                     c = lc;
                     e++;
                  }
               }
               const int idx = idx0 - cp.koff[dd][e][c];
               const int ncc = d_nco(c), ncc1 = d_nco(c + 1);
               const int ie = idx / (ncc * nd);
               const int rem = idx - ie * ncc * nd;
               const int ic = rem / nd, id = rem - ic * nd;
               const int dir = t_first[dd * 10 + id];
               const int idm = t_minus[(dd * 10 + id) * 3 + dir];
               const int icp = t_plus[(c * 28 + ic) * 3 + dir];
               const double v1 = in[cp.koff[dd - 1][e][c + 1] + (ie * ncc1 + icp) * ndm + idm];
               const double v0 = in[cp.koff[dd - 1][e][c] + (ie * ncc + ic) * ndm + idm];
               WS[l * lws + dst * cap + idx0] = v1 + CD[dir] * v0;
            }
            __syncthreads();
            srcb = dst;
         }
         // bra HRR (b -> a): layer bb from layer bb-1, blocks a in [la, LA-bb]
         for (int bb = 1; bb <= lb; bb++) {
            const int dst = (srcb + 1) % 3;
            const int nbb = d_nco(bb), nbm = d_nco(bb - 1);
            const int sz = cp.bsize[bb];
            for (int idx1 = tid; idx1 < nb * sz; idx1 += NT) {
               const int l = fdiv(idx1, sz), idx0 = idx1 - l * sz;
               const double *in = WS + l * lws + srcb * cap;
               int a = la;
               while (idx0 >= cp.boff[bb][a] + d_nco(a) * nbb * ncd) a++;
               const int idx = idx0 - cp.boff[bb][a];
               const int ia = idx / (nbb * ncd);
               const int rem = idx - ia * nbb * ncd;
               const int ib = rem / ncd, icd = rem - ib * ncd;
               const int dir = t_first[bb * 10 + ib];
               const int ibm = t_minus[(bb * 10 + ib) * 3 + dir];
               const int iap = t_plus[(a * 28 + ia) * 3 + dir];
               const double v1 = in[cp.boff[bb - 1][a + 1] + (iap * nbm + ibm) * ncd + icd];
               const double v0 = in[cp.boff[bb - 1][a] + (ia * nbm + ibm) * ncd + icd];
               WS[l * lws + dst * cap + idx0] = v1 + AB[dir] * v0;
            }
            __syncthreads();
            srcb = dst;
         }
         // the first block of the last layer is (la lb|lc ld)
         for (int idx = tid; idx < nb * Nc; idx += NT) {
            const int l = fdiv(idx, Nc), i = idx - l * Nc;
            WS[l * lws + 3 * cap + i] += WS[l * lws + srcb * cap + i];
         }
         __syncthreads();
      }
      // cartesian -> spherical: four stages, each moves the leading index last
      {
         const int ls[4] = {la, lb, lc, ld};
         int dims[4] = {d_nco(la), d_nco(lb), d_nco(lc), d_nco(ld)};
         int in_off = 3 * cap;
         for (int s = 0; s < 4; s++) {
            const int out_off = (s & 1) * cap;
            const int k = dims[0];
            const int n = dims[1] * dims[2] * dims[3];
            const int m = 2 * ls[s] + 1;
            const double *c2s = (const double *)(tabs + CQ_TAB_C2S) + c_c2s_off[ls[s]];
            for (int idx1 = tid; idx1 < nb * n * m; idx1 += NT) {
               const int l = fdiv(idx1, n * m), idx = idx1 - l * n * m;
               const int j = idx / m, i = idx - j * m;
               const double *in = WS + l * lws + in_off;
               double v = 0.0;
               for (int qq = 0; qq < k; qq++) v += c2s[i * k + qq] * in[qq * n + j];
               WS[l * lws + out_off + idx] = v;
            }
            __syncthreads();
            in_off = out_off;
            const int nd0 = dims[1], nd1 = dims[2], nd2 = dims[3];
            dims[0] = nd0;
            dims[1] = nd1;
            dims[2] = nd2;
            dims[3] = m;
         }
      }
      // spherical integrals of contraction l0+l at WS + l lws + cap
      if (reduce) {
         for (int idx = tid; idx < nb * Ns; idx += NT) {
            const int l = fdiv(idx, Ns);
            Iall[(l0 + l) * Ns + (idx - l * Ns)] = WS[l * lws + cap + (idx - l * Ns)];
         }
         continue;
      }
      if (out_mode) {
         // the integrals themselves, in the reference's TRA order
         double *O = kt.out + ((size_t)ks.base[0] | (size_t)ks.base[1] << 32);
         for (int idx = tid; idx < nb * Ns; idx += NT) {
            const int l = fdiv(idx, Ns), tt = idx - l * Ns;
            const int q = l0 + l;
            const int ia = q / (nlb * nlc * nld), ib = (q / (nlc * nld)) % nlb, ic = (q / nld) % nlc, id = q % nld;
            const int isa = tt / nsbcd, isb = (tt / nscd) % nsb, isc = (tt / nsd) % nsc, isd = tt % nsd;
            O[tra_index(ia, isa, ib, isb, ic, isc, id, isd, nlb, nlc, nld, nsa, nsb, nsc, nsd)] = WS[l * lws + cap + tt];
         }
         continue;
      }
      if (use_acc) {
         // the batch's contractions into the K accumulators
         kacc_add_pass<NT>(kt, acc, y.kacc, ks, l0, nb, nlb, nlc, nld, nsa, nsb, nsc, nsd, NA, NB, NC, ND,
                           [&](int l, int isa, int isb, int isc, int isd) {
                              return WS[l * lws + cap + ((isa * nsb + isb) * nsc + isc) * nsd + isd];
                           });
         continue;
      }
      // one contraction at a time: 4 atomic updates per spherical integral
      for (int idx = tid; idx < nb * Ns; idx += NT) {
         const int l = fdiv(idx, Ns), tt = idx - l * Ns;
         size_t off[4];
         uint32_t ld01 = ks.ld01, ld23 = ks.ld23, flags = ks.flags;
         if (expl) {
            const KsRec x = b.ksx[ks.expl + l0 + l];
            for (int q = 0; q < 4; q++) off[q] = x.off[q];
            ld01 = x.ld01;
            ld23 = x.ld23;
            flags = x.flags;
         } else {
            ks_offsets(ks, nlb, nlc, nld, l0 + l, off);
         }
         const double fac = symm[(flags >> 4) & 7] * kt.hf_fac;
         const unsigned ldac = ld01 & 0xffff, ldad = ld01 >> 16, ldbc = ld23 & 0xffff, ldbd = ld23 >> 16;
         const unsigned isa = tt / nsbcd, isb = (tt / nscd) % nsb, isc = (tt / nsd) % nsc, isd = tt % nsd;
         const size_t Iac = off[0] + blk_idx(flags & 1, ldac, isa, isc);
         const size_t Iad = off[1] + blk_idx(flags & 2, ldad, isa, isd);
         const size_t Ibc = off[2] + blk_idx(flags & 4, ldbc, isb, isc);
         const size_t Ibd = off[3] + blk_idx(flags & 8, ldbd, isb, isd);
         const double iabcd = -fac * WS[l * lws + cap + tt];
         for (int sp = 0; sp < kt.nspin; sp++) {
            const double *P = kt.P[sp];
            double *K = kt.K[sp];
            unsafeAtomicAdd(&K[Ibd], iabcd * P[Iac]);
            unsafeAtomicAdd(&K[Ibc], iabcd * P[Iad]);
            unsafeAtomicAdd(&K[Iad], iabcd * P[Ibc]);
            unsafeAtomicAdd(&K[Iac], iabcd * P[Ibd]);
         }
      }
   }
   if (use_acc) {
      __syncthreads();
      kacc_flush<NT>(kt, ks, acc, y.kacc, nsa, nsb, nsc, nsd, NA, NB, NC, ND);
      return;
   }
   if (!reduce) return;
   __syncthreads();
   // the same updates, summed over the set quartet's contractions first:
   // K_bd += i P_ac, K_bc += i P_ad, K_ad += i P_bc, K_ac += i P_bd
   const double fac = symm[(ks.flags >> 4) & 7] * kt.hf_fac;
   if (fac == 0.0) return;
   const unsigned ldac = ks.ld01 & 0xffff, ldad = ks.ld01 >> 16, ldbc = ks.ld23 & 0xffff, ldbd = ks.ld23 >> 16;
   const bool Tac = ks.flags & 1, Tad = ks.flags & 2, Tbc = ks.flags & 4, Tbd = ks.flags & 8;
   auto Iat = [&](int ia, int isa, int ib, int isb, int ic, int isc, int id, int isd) {
      return Iall[(((ia * nlb + ib) * nlc + ic) * nld + id) * Ns + ((isa * nsb + isb) * nsc + isc) * nsd + isd];
   };
   auto offs = [&](int ia, int ib, int ic, int id, int q) {
      switch (q) {
         case 0: return (long long)ks.base[0] + (long long)ia * ks.step[0] + (long long)ic * ks.step[1];
         case 1: return (long long)ks.base[1] + (long long)ia * ks.step[2] + (long long)id * ks.step[3];
         case 2: return (long long)ks.base[2] + (long long)ib * ks.step[4] + (long long)ic * ks.step[5];
         default: return (long long)ks.base[3] + (long long)ib * ks.step[6] + (long long)id * ks.step[7];
      }
   };
   for (int sp = 0; sp < kt.nspin; sp++) {
      const double *P = kt.P[sp];
      double *K = kt.K[sp];
      for (int x = tid; x < NB * ND; x += NT) {  // K_bd
         const int ib = (x / ND) / nsb, isb = (x / ND) % nsb, id = (x % ND) / nsd, isd = (x % ND) % nsd;
         double sum = 0.0;
         for (int ia = 0; ia < nla; ia++)
            for (int ic = 0; ic < nlc; ic++) {
               const size_t o = (size_t)offs(ia, 0, ic, 0, 0);
               for (int isa = 0; isa < nsa; isa++)
                  for (int isc = 0; isc < nsc; isc++)
                     sum += Iat(ia, isa, ib, isb, ic, isc, id, isd) * P[o + blk_idx(Tac, ldac, isa, isc)];
            }
         unsafeAtomicAdd(&K[(size_t)offs(0, ib, 0, id, 3) + blk_idx(Tbd, ldbd, isb, isd)], -fac * sum);
      }
      for (int x = tid; x < NB * NC; x += NT) {  // K_bc
         const int ib = (x / NC) / nsb, isb = (x / NC) % nsb, ic = (x % NC) / nsc, isc = (x % NC) % nsc;
         double sum = 0.0;
         for (int ia = 0; ia < nla; ia++)
            for (int id = 0; id < nld; id++) {
               const size_t o = (size_t)offs(ia, 0, 0, id, 1);
               for (int isa = 0; isa < nsa; isa++)
                  for (int isd = 0; isd < nsd; isd++)
                     sum += Iat(ia, isa, ib, isb, ic, isc, id, isd) * P[o + blk_idx(Tad, ldad, isa, isd)];
            }
         unsafeAtomicAdd(&K[(size_t)offs(0, ib, ic, 0, 2) + blk_idx(Tbc, ldbc, isb, isc)], -fac * sum);
      }
      for (int x = tid; x < NA * ND; x += NT) {  // K_ad
         const int ia = (x / ND) / nsa, isa = (x / ND) % nsa, id = (x % ND) / nsd, isd = (x % ND) % nsd;
         double sum = 0.0;
         for (int ib = 0; ib < nlb; ib++)
            for (int ic = 0; ic < nlc; ic++) {
               const size_t o = (size_t)offs(0, ib, ic, 0, 2);
               for (int isb = 0; isb < nsb; isb++)
                  for (int isc = 0; isc < nsc; isc++)
                     sum += Iat(ia, isa, ib, isb, ic, isc, id, isd) * P[o + blk_idx(Tbc, ldbc, isb, isc)];
            }
         unsafeAtomicAdd(&K[(size_t)offs(ia, 0, 0, id, 1) + blk_idx(Tad, ldad, isa, isd)], -fac * sum);
      }
      for (int x = tid; x < NA * NC; x += NT) {  // K_ac
         const int ia = (x / NC) / nsa, isa = (x / NC) % nsa, ic = (x % NC) / nsc, isc = (x % NC) % nsc;
         double sum = 0.0;
         for (int ib = 0; ib < nlb; ib++)
            for (int id = 0; id < nld; id++) {
               const size_t o = (size_t)offs(0, ib, 0, id, 3);
               for (int isb = 0; isb < nsb; isb++)
                  for (int isd = 0; isd < nsd; isd++)
                     sum += Iat(ia, isa, ib, isb, ic, isc, id, isd) * P[o + blk_idx(Tbd, ldbd, isb, isd)];
            }
         unsafeAtomicAdd(&K[(size_t)offs(ia, 0, ic, 0, 0) + blk_idx(Tac, ldac, isa, isc)], -fac * sum);
      }
   }
}

// ---------------------------------------------------------------------------
// Kronecker path of k_gqs for classes whose HRR layers do not fit in LDS
// and whose contraction ran on the matrix cores.  The HRR of a side,
// (x, y + 1_d| = (x + 1_d, y| + R_d (x y| with R = AB (bra) or CD (ket)
// (reference src/compute_HRR.cpp), unrolls to
//    (x y| = sum_{0 <= i <= y} prod_d C(y_d, i_d) R_d^(y_d - i_d) (x + i, 0|,
// so with the spherical transform (src/compute_SPH.cpp:58-237) folded in,
// the spherical pair (s1 s2) of a side is a fixed combination of the side's
// block components g (shells l1 .. l1 + l2, shell outer):
//    M[s1 s2][g] = sum_x c2s(s1, x) T[s2][g - x],
//    T[s2][i]    = sum_{y >= i} c2s(s2, y) prod_d C(y_d, i_d) R_d^(y_d - i_d),
// i running over the vectors with |i| <= l2 (index ioff3(|i|) + cartesian
// index).  A row's spherical integrals are then M_bra C M_ket^T, C its
// contracted blocks (bra components x ket components): two small dense
// products per contraction and row on the matrix cores, summed over the
// set quartet's rows in the accumulators.
// ---------------------------------------------------------------------------
typedef double kdbl4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64, lane layout as in the contraction kernel: lane l
// holds A[l % 16][l / 16] and B[l / 16][l % 16]; element r of its D is
// D[4 r + l / 16][l % 16]
__device__ __forceinline__ kdbl4 kron_mfma(double a, double b, kdbl4 c)
{
   return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// binomial coefficient C(n, k), 0 <= k <= n <= 3
__device__ __forceinline__ double binom3(int n, int k)
{
   return (n == 3 && (k == 1 || k == 2)) ? 3.0 : ((n == 2 && k == 1) ? 2.0 : 1.0);
}

__device__ __forceinline__ double ipow3(double r, int n)
{
   return n == 0 ? 1.0 : (n == 1 ? r : (n == 2 ? r * r : r * r * r));
}

// cartesian components of the shells below k
__host__ __device__ __forceinline__ int ioff3(int k) { return k * (k + 1) * (k + 2) / 6; }

// T[s2][iv] of a side whose second shell is l2
__device__ __forceinline__ double kron_t(int l2, int s2, int iv, const double R[3], const double *c2)
{
   int K = 0;
   while (iv >= ioff3(K + 1)) K++;
   const int ci = iv - ioff3(K);
   const int i0 = c_cart.comp[K][ci][0], i1 = c_cart.comp[K][ci][1], i2 = c_cart.comp[K][ci][2];
   const int n2 = d_nco(l2);
   double v = 0.0;
   for (int y = 0; y < n2; y++) {
      const int y0 = c_cart.comp[l2][y][0], y1 = c_cart.comp[l2][y][1], y2 = c_cart.comp[l2][y][2];
      if (y0 < i0 || y1 < i1 || y2 < i2) continue;
      v += c2[s2 * n2 + y] * binom3(y0, i0) * binom3(y1, i1) * binom3(y2, i2) * ipow3(R[0], y0 - i0) *
           ipow3(R[1], y1 - i1) * ipow3(R[2], y2 - i2);
   }
   return v;
}

// M[s][g] of a side (shells l1, l2) from its T table (ni columns)
__device__ __forceinline__ double kron_m(int l1, int l2, int s, int g, const double *T, int ni, const double *c1)
{
   const int ns2 = 2 * l2 + 1;
   const int s1 = s / ns2, s2 = s - s1 * ns2;
   int f = l1, gg = g;
   while (gg >= d_nco(f)) {
      gg -= d_nco(f);
      f++;
   }
   const int g0 = c_cart.comp[f][gg][0], g1 = c_cart.comp[f][gg][1], g2 = c_cart.comp[f][gg][2];
   const int K = f - l1, n1 = d_nco(l1);
   const double *Ts = T + s2 * ni + ioff3(K);
   double v = 0.0;
   for (int x = 0; x < n1; x++) {
      const int i0 = g0 - c_cart.comp[l1][x][0], i1 = g1 - c_cart.comp[l1][x][1], i2 = g2 - c_cart.comp[l1][x][2];
      if (i0 < 0 || i1 < 0 || i2 < 0) continue;
      v += c1[s1 * n1 + x] * Ts[(K - i0) * (K - i0 + 1) / 2 + i2];
   }
   return v;
}

// One (set quartet, class) on the Kronecker path, four wavefronts, kb =
// y.kron_kb contractions per pass.  LDS at y.o_kron: kb blocks of YB
// doubles (first Y = C M_ket^T for the pass's contractions stacked, bra
// components x ket spherical pairs; then the contractions' integrals, one
// block each; the second product takes the pass's contractions side by
// side), then the two T tables; at y.o_kc (own contraction) the pass's contracted
// blocks and coefficient products.  W (the workgroup's HBM slice, or LDS at
// y.o_kmap) holds the maps of up to kron_ws / (one row's maps) rows: all
// rows of the set quartet when they fit (built once), else one row at a
// time (rebuilt per pass).  The pass's integral tiles (at most 16: four per
// wavefront) stay in the wavefronts' accumulators over the rows; then they
// go to LDS and to the K updates (or the integral output) of the pass's
// contractions.
template <int NT>
__device__ void cq_kron(const ClassPlan &cp, int cls_k, const TypePlan &tp, const DevTables &t, const BatchDev &b,
                        const KTargets &kt, const CqLayout &y, double *W, uint32_t gi, const char *tabs,
                        const double *co, const int *__restrict__ srcmap)
{
   static_assert(NT == 256, "four wavefronts");
   extern __shared__ double lds[];
   const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
   const int la = cp.la, lb = cp.lb, lc = cp.lc, ld = cp.ld;
   const int ncp = cp.ncp;
   const int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1, nsd = 2 * ld + 1;
   const int nsab = nsa * nsb, nscd = nsc * nsd, nsbcd = nsb * nscd;
   int nbra = 0, nket = 0;
   for (int e = la; e <= cp.LA; e++) nbra += d_nco(e);
   for (int f = lc; f <= cp.LC; f++) nket += d_nco(f);
   const int TM1 = (nbra + 15) >> 4, TN = (nscd + 15) >> 4, TM2 = (nsab + 15) >> 4;
   const int YS = 16 * TN, YB = 16 * max(TM1, TM2) * YS;
   const int KS1 = (nket + 3) >> 2, KS2 = (nbra + 3) >> 2;
   const int nik = ioff3(ld + 1), nib = ioff3(lb + 1);
   const int KB = y.kron_kb;
   double *Y = lds + y.o_kron;
   double *Tk = Y + KB * YB, *Tb = Tk + nsd * nik;
   const int MS = nscd * nket + nsab * nbra;
   const int rc = y.kron_ws / MS;
   const double *c2s = (const double *)(tabs + CQ_TAB_C2S);
   const double *ca = c2s + c_c2s_off[la], *cb = c2s + c_c2s_off[lb], *cc = c2s + c_c2s_off[lc],
                *cd = c2s + c_c2s_off[ld];
   const int sh = tp.cls_sh[cls_k];
   const int nla = tp.nl[0][sh & 15], nlb = tp.nl[1][(sh >> 4) & 15], nlc = tp.nl[2][(sh >> 8) & 15],
             nld = tp.nl[3][(sh >> 12) & 15];
   const GroupRec &g = b.groups[gi];
   const KsGroup ks = group_ks(b, tp, g, cls_k);
   const int nq = (int)(ks.flags >> 16);
   if (nq == 0) return;
   const bool expl = (ks.flags & KS_EXPLICIT) != 0;
   const bool out_mode = (ks.flags & KS_OUTPUT) != 0;
   const uint32_t nr = g.row1 - g.row0;
   const bool cached = nr <= (uint32_t)rc;
   // K accumulators in LDS (summed over the contractions before the atomics)
   const bool use_acc = y.o_kacc >= 0 && !expl && !out_mode && nq == nla * nlb * nlc * nld &&
                        kt.nspin * y.kacc <= y.kacc_cap;
   const int NA = nla * nsa, NB = nlb * nsb, NC = nlc * nsc, ND = nld * nsd;
   double *kap = lds + y.o_kacc;
   if (use_acc)
      for (int i = tid; i < kt.nspin * y.kacc; i += NT) kap[i] = 0.0;

   // the two maps of row r into slot m of W
   auto build = [&](uint32_t r, int m) {
      double AB[3], CD[3];
      row_ab_cd(g, b.rows[r], b.data, t, AB, CD);
      __syncthreads();   // the T tables and the slot are no longer read
      for (int idx = tid; idx < nsd * nik + nsb * nib; idx += NT) {
         if (idx < nsd * nik) {
            Tk[idx] = kron_t(ld, idx / nik, idx % nik, CD, cd);
         } else {
            const int j = idx - nsd * nik;
            Tb[j] = kron_t(lb, j / nib, j % nib, AB, cb);
         }
      }
      __syncthreads();
      double *Mk = W + (size_t)m * MS, *Mb = Mk + nscd * nket;
      for (int idx = tid; idx < MS; idx += NT) {
         if (idx < nscd * nket) {
            Mk[idx] = kron_m(lc, ld, idx / nket, idx % nket, Tk, nik, cc);
         } else {
            const int j = idx - nscd * nket;
            Mb[j] = kron_m(la, lb, j / nbra, j % nbra, Tb, nib, ca);
         }
      }
   };
   if (cached)
      for (uint32_t r = g.row0; r < g.row1; r++) build(r, (int)(r - g.row0));

   for (int l0 = 0; l0 < nq; l0 += KB) {
      const int nb = min(KB, nq - l0);
      const int TNS = (nb * nscd + 15) >> 4;   // integral tile columns of the pass
      kdbl4 acc[4];
#pragma unroll
      for (int i = 0; i < 4; i++) acc[i] = kdbl4{0.0, 0.0, 0.0, 0.0};
      for (uint32_t r = g.row0; r < g.row1; r++) {
         if (!cached) build(r, 0);
         __syncthreads();   // the maps are written, the previous Y is no longer read
         const double *Mk = W + (size_t)(cached ? r - g.row0 : 0) * MS, *Mb = Mk + nscd * nket;
         const double *Cb;   // contraction l0 + l of the row at Cb + l ncp
         if (y.eco) {
            // contracted on the matrix cores (k_eco): the pass's blocks are
            // contiguous in the contraction buffer; where the LDS has room
            // (o_kc) they are copied there with coalesced loads, once per
            // pass, instead of each tile lane loading its own 32 bytes
            Cb = b.eco + (size_t)(r - b.r0) * b.eco_row + y.eco_off + (size_t)l0 * ncp;
            if (y.o_kc >= 0) {
               double *Cs = lds + y.o_kc;
               for (int j = tid; j < nb * ncp; j += NT) Cs[j] = Cb[j];
               __syncthreads();
               Cb = Cs;
            }
         } else {
            // the class's own contraction of the row's primitive blocks for
            // the pass's contractions, into LDS (as the layered path's)
            double *Cs = lds + y.o_kc, *Kp = Cs + KB * ncp;
            const RowRec row = b.rows[r];
            const uint32_t *IP = b.prm_ip + row.prm_off;
            const double *A0 = b.ac0 + (size_t)(row.prm_off - b.p0) * tp.ncp;
            for (int j = tid; j < nb * ncp; j += NT) Cs[j] = 0.0;
            for (int pc = 0; pc < (int)row.n_prm; pc += CQ_PCH) {
               const int np = min(CQ_PCH, (int)row.n_prm - pc);
               __syncthreads();   // the previous step's products are no longer read
               for (int i = tid; i < nb * CQ_PCH; i += NT) {
                  const int l = i / CQ_PCH, pp = i - l * CQ_PCH;
                  Kp[i] = pp < np ? kprod_lds(tp, co, cls_k, l0 + l, IP[pc + pp]) : 0.0;
               }
               __syncthreads();
               const double *A = A0 + (size_t)pc * tp.ncp;
               for (int j = tid; j < ncp; j += NT) {
                  const int src = srcmap[j];
                  double av[CQ_PCH];
#pragma unroll
                  for (int pp = 0; pp < CQ_PCH; pp++) av[pp] = pp < np ? A[(size_t)pp * tp.ncp + src] : 0.0;
                  for (int l = 0; l < nb; l++) {
                     double sum = 0.0;
#pragma unroll
                     for (int pp = 0; pp < CQ_PCH; pp++) sum += Kp[l * CQ_PCH + pp] * av[pp];
                     Cs[l * ncp + j] += sum;
                  }
               }
            }
            __syncthreads();
            Cb = Cs;
         }
         // Y = [C_l0; C_l1; ...] M_ket^T: the pass's contractions stacked
         // (row m = l nbra + bc: whole 16-row tiles over all of them, not
         // per contraction), its TMS x TN tiles over the wavefronts
         const int TMS = (nb * nbra + 15) >> 4;
         for (int tile = wv; tile < TMS * TN; tile += 4) {
            const int tm = tile / TN, tn = tile - tm * TN;
            const int m = 16 * tm + (lane & 15), s = 16 * tn + (lane & 15);
            const bool mv = m < nb * nbra, sv = s < nscd;
            const int l = mv ? m / nbra : 0, bc = mv ? m - l * nbra : 0;
            const double *Cl = Cb + (size_t)l * ncp;
            int e = la, ie = bc;
            while (e < cp.LA && ie >= d_nco(e)) {
               ie -= d_nco(e);
               e++;
            }
            // lane's ket component 4 k4 + lane / 16 as (shell f, index jf)
            int f = lc, jf = lane >> 4;
            while (f < cp.LC && jf >= d_nco(f)) {
               jf -= d_nco(f);
               f++;
            }
            const double *Mks = Mk + s * nket;
            kdbl4 d = {0.0, 0.0, 0.0, 0.0};
            for (int gk = lane >> 4; gk < 4 * KS1; gk += 4) {
               const bool gv = gk < nket;
               const double av = (mv && gv) ? Cl[cp.cpoff[e][f] + ie * d_nco(f) + jf] : 0.0;
               const double bm = (gv && sv) ? Mks[gk] : 0.0;
               d = kron_mfma(av, bm, d);
               jf += 4;
               while (f < cp.LC && jf >= d_nco(f)) {
                  jf -= d_nco(f);
                  f++;
               }
            }
#pragma unroll
            for (int rr = 0; rr < 4; rr++) Y[(16 * tm + 4 * rr + (lane >> 4)) * YS + 16 * tn + (lane & 15)] = d[rr];
         }
         __syncthreads();
         // integrals_l += M_bra Y_l for the pass's contractions side by side
         // (column n = l nscd + scd: whole 16-column tiles over all of them),
         // wavefront wv the tiles wv, wv + 4, ... (at most 4)
#pragma unroll
         for (int i = 0; i < 4; i++) {
            const int tile = wv + 4 * i;
            if (tile < TM2 * TNS) {
               const int tm = tile / TNS, tn = tile - tm * TNS;
               const int n = 16 * tn + (lane & 15);
               const bool nv = n < nb * nscd;
               const int l = nv ? n / nscd : 0, scd = nv ? n - l * nscd : 0;
               const double *Yl = Y + (size_t)l * nbra * YS + scd;
               const int sab = 16 * tm + (lane & 15);
               for (int k4 = 0; k4 < KS2; k4++) {
                  const int bk = 4 * k4 + (lane >> 4);
                  // (rows past nbra belong to the next contraction or are unwritten)
                  const bool kv = bk < nbra;
                  const double av = (sab < nsab && kv) ? Mb[sab * nbra + bk] : 0.0;
                  const double yv = (kv && nv) ? Yl[bk * YS] : 0.0;
                  acc[i] = kron_mfma(av, yv, acc[i]);
               }
            }
         }
      }
      // the pass's integrals into LDS (over Y): contraction l at Y + l YB, [bra pair][YS]
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; i++) {
         const int tile = wv + 4 * i;
         if (tile < TM2 * TNS) {
            const int tm = tile / TNS, tn = tile - tm * TNS;
            const int n = 16 * tn + (lane & 15);
            if (n < nb * nscd) {
               const int l = n / nscd, scd = n - l * nscd;
#pragma unroll
               for (int rr = 0; rr < 4; rr++) Y[l * YB + (16 * tm + 4 * rr + (lane >> 4)) * YS + scd] = acc[i][rr];
            }
         }
      }
      __syncthreads();
      // their K updates (or their integrals, in the TRA order)
      if (use_acc) {
         kacc_add_pass<NT>(kt, kap, y.kacc, ks, l0, nb, nlb, nlc, nld, nsa, nsb, nsc, nsd, NA, NB, NC, ND,
                           [&](int l, int isa, int isb, int isc, int isd) {
                              return Y[l * YB + (isa * nsb + isb) * YS + isc * nsd + isd];
                           });
         continue;
      }
      for (int l = 0; l < nb; l++) {
         const int q = l0 + l;
         const double *Il = Y + l * YB;
         const int ia = q / (nlb * nlc * nld), ib = (q / (nlc * nld)) % nlb, ic = (q / nld) % nlc, id = q % nld;
         size_t off[4] = {0, 0, 0, 0};
         uint32_t ld01 = ks.ld01, ld23 = ks.ld23, flags = ks.flags;
         if (expl) {
            const KsRec x = b.ksx[ks.expl + q];
            for (int qq = 0; qq < 4; qq++) off[qq] = x.off[qq];
            ld01 = x.ld01;
            ld23 = x.ld23;
            flags = x.flags;
         } else if (!out_mode) {
            ks_offsets(ks, nlb, nlc, nld, q, off);
         }
         const double fac = symm_code_fac((flags >> 4) & 7) * kt.hf_fac;
         const unsigned ldac = ld01 & 0xffff, ldad = ld01 >> 16, ldbc = ld23 & 0xffff, ldbd = ld23 >> 16;
         for (int tt = tid; tt < cp.Ns; tt += NT) {
            const unsigned isa = tt / nsbcd, isb = (tt / nscd) % nsb, isc = (tt / nsd) % nsc, isd = tt % nsd;
            const double v = Il[(tt / nscd) * YS + tt % nscd];
            if (out_mode) {
               double *O = kt.out + ((size_t)ks.base[0] | (size_t)ks.base[1] << 32);
               O[tra_index(ia, isa, ib, isb, ic, isc, id, isd, nlb, nlc, nld, nsa, nsb, nsc, nsd)] = v;
               continue;
            }
            const size_t Iac = off[0] + blk_idx(flags & 1, ldac, isa, isc);
            const size_t Iad = off[1] + blk_idx(flags & 2, ldad, isa, isd);
            const size_t Ibc = off[2] + blk_idx(flags & 4, ldbc, isb, isc);
            const size_t Ibd = off[3] + blk_idx(flags & 8, ldbd, isb, isd);
            const double iabcd = -fac * v;
            for (int sp = 0; sp < kt.nspin; sp++) {
               const double *P = kt.P[sp];
               double *K = kt.K[sp];
               unsafeAtomicAdd(&K[Ibd], iabcd * P[Iac]);
               unsafeAtomicAdd(&K[Ibc], iabcd * P[Iad]);
               unsafeAtomicAdd(&K[Iad], iabcd * P[Ibc]);
               unsafeAtomicAdd(&K[Iac], iabcd * P[Ibd]);
            }
         }
      }
   }
   if (use_acc) {
      __syncthreads();
      kacc_flush<NT>(kt, ks, kap, y.kacc, nsa, nsb, nsc, nsd, NA, NB, NC, ND);
   }
}

// Work items are (set quartet, class of this part) pairs, the classes of a
// set quartet on consecutive items: a type with few, large set quartets (an
// spdf centre's 16^4 primitive quartets and 256 classes) still fills the
// chip, and the set quartet's blocks are read by neighbouring workgroups
// while they sit in L2.  WPE wavefronts per SIMD: at most 128 VGPRs for 4
// (the kernel waits on memory, and more wavefronts in flight outweigh the
// few values it then keeps in scratch); 3 or 2 (168 or 256 VGPRs) where the
// workgroup's LDS leaves room for only that many workgroups per CU anyway
template <int NT, int WPE = 4>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_gqs(const ClassPlan *__restrict__ cps, const CqLayout *__restrict__ lays,
                                                    const int *__restrict__ srcmaps, const TypePlan *__restrict__ tpp,
                                                    DevTables t, BatchDev b, KTargets kt, double *gws, int gws_stride,
                                                    int cls_bytes, int part)
{
   const TypePlan &tp = *tpp;
   const int tid = threadIdx.x;
   // LDS: [class workspace, cls_bytes][class plan][coefficients nco_tot][primitive indices CQ_PCH][tables][classes of the part]
   ClassPlan *cpl = (ClassPlan *)((char *)lds_base() + cls_bytes);
   double *co = (double *)((char *)cpl + CPL_BYTES);
   uint32_t *ipb = (uint32_t *)(co + tp.nco_tot);
   char *tabs = (char *)lds_base() + ((cls_bytes + CPL_BYTES + 8 * tp.nco_tot + 4 * CQ_PCH + 15) & ~15);
   int *plist = (int *)(tabs + CQ_TAB_BYTES);   // [MAXTC + 1]: count, then the part's classes
   for (int i = tid; i < GINT_C2S_SIZE; i += NT) ((double *)(tabs + CQ_TAB_C2S))[i] = c_c2s[i];
   for (int i = tid; i < 40; i += NT) ((unsigned char *)(tabs + CQ_TAB_FIRST))[i] = c_cart.first[i / 10][i % 10];
   for (int i = tid; i < 120; i += NT)
      ((signed char *)(tabs + CQ_TAB_MINUS))[i] = c_cart.minus[i / 30][(i / 3) % 10][i % 3];
   for (int i = tid; i < 588; i += NT) ((signed char *)(tabs + CQ_TAB_PLUS))[i] = c_cart.plus[i / 84][(i / 3) % 28][i % 3];
   if (tid == 0) {
      int n = 0;
      for (int k = 0; k < tp.nclass; k++)
         if (!lays[k].skip && lays[k].part == part) plist[1 + n++] = k;
      plist[0] = n;
   }
   __syncthreads();
   const uint32_t npc = (uint32_t)plist[0];
   const uint32_t nitems = (b.g1 - b.g0) * npc;
   uint32_t cur_g = 0xffffffffu;
   for (uint32_t item = blockIdx.x; item < nitems; item += gridDim.x) {
      const uint32_t gi = b.g0 + item / npc;
      const int k = plist[1 + item % npc];
      const GroupRec &g = b.groups[gi];
      const bool setlevel = (g.ksmode & KS_SETLEVEL) != 0;
      if (!setlevel && ks_count(b.ksg[(size_t)g.ks + k].flags) == 0) continue;
      __syncthreads();
      if (gi != cur_g) {
         // the set quartet's contraction coefficients, shell by shell
         for (int tc = 0; tc < 4; tc++) {
            const int npt = ((g.np4 >> (4 * tc)) & 15) + 1;
            for (int s = 0; s < tp.nsh[tc]; s++) {
               const int n = tp.nl[tc][s] * MAXNP;
               for (int i = tid; i < n; i += NT) {
                  const int inl = i / MAXNP, ip = i % MAXNP;
                  co[tp.co_off[tc][s] + i] = ip < npt ? b.data[g.Kof[tc][s] + inl * npt + ip] : 0.0;
               }
            }
         }
         cur_g = gi;
      }
      const CqLayout &y = lays[k];
      {
         const int *src = (const int *)(cps + k);
         int *dst = (int *)cpl;
         for (int i = tid; i < (int)(sizeof(ClassPlan) / 4); i += NT) dst[i] = src[i];
      }
      __syncthreads();
      double *WS = y.ws_global ? gws + (size_t)blockIdx.x * gws_stride : lds_base() + y.o_ws;
      // the Kronecker path where the class's schedule takes it (four
      // wavefronts, its blocks contracted on the matrix cores or by itself)
      // and the chunk has no class-dependent screening to add; else the
      // layered one
      bool kron = false;
      if constexpr (NT == 256)
         kron = y.kron && y.kron_ws && (y.eco || y.o_kc >= 0) && min(b.fix_count[0], b.fix_cap) == 0;
      if (kron)
         cq_kron<256>(*cpl, k, tp, t, b, kt, y, y.o_kmap >= 0 ? lds_base() + y.o_kmap : WS, gi, tabs, co,
                      srcmaps + y.o_src);
      else
         cq_one<NT>(*cpl, k, tp, t, b, kt, y, WS, srcmaps + y.o_src, gi, co, ipb, tabs);
   }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static constexpr size_t LDS_LIMIT = 64 * 1024;

size_t fm_r_lds_bytes(const TypePlan &tp)
{
   const FmLayout y = fm_layout(tp, false);
   return (size_t)y.bytes <= LDS_LIMIT ? (size_t)y.bytes : 0;
}

int launch_fm_r(const TypePlan &tp, const DevTables &t, const BatchDev &b, void *stream)
{
   if (b.p1 <= b.p0) return 0;
   const uint32_t n = b.p1 - b.p0;
   // two primitive quartets per wavefront when each takes at most half of
   // the lanes' cells and its cell sum fits 4 values per lane
   const FmLayout yp = fm_layout(tp, false, true);
   if (t.Ng <= 32 && tp.nh <= 256 && (size_t)yp.bytes <= LDS_LIMIT) {
      // workgroups stay resident and stride over the pairs
      const unsigned grid = std::min((n + 1) / 2, 16384u);
      hipLaunchKernelGGL((k_fm_r<false, true>), dim3(grid), dim3(64), yp.bytes, (hipStream_t)stream, b.tp, t, b, yp);
      return hipGetLastError() == hipSuccess ? 0 : -1;
   }
   const FmLayout y = fm_layout(tp, false);
   if ((size_t)y.bytes > LDS_LIMIT) return -3;
   const unsigned grid = std::min(n, 16384u);
   hipLaunchKernelGGL((k_fm_r<false, false>), dim3(grid), dim3(64), y.bytes, (hipStream_t)stream, b.tp, t, b, y);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fixup(const TypePlan &tp, const DevTables &t, const BatchDev &b, uint32_t n_fix, void *stream)
{
   const FmLayout y = fm_layout(tp, true);
   if ((size_t)y.bytes > LDS_LIMIT) return -3;
   const unsigned grid = std::max(1u, std::min(n_fix, 4096u));
   hipLaunchKernelGGL((k_fm_r<true, false>), dim3(grid), dim3(64), y.bytes, (hipStream_t)stream, b.tp, t, b, y);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

// LDS plan of one class in k_gqs: as many contractions per pass as fit in
// `budget` doubles, the workspace in HBM when not even one does
static CqLayout cq_layout(const ClassPlan &cp, int nq, int budget_all, int kacc)
{
   CqLayout y;
   y.lws = 3 * cp.ycap + cp.Nc;
   const int need_I = nq * cp.Ns;
   y.reduce = need_I <= 2048 ? 1 : 0;
   // K accumulators where the integrals of all contractions do not fit: the
   // four K blocks of the set quartet (kacc doubles per spin), within half
   // the budget, for two spins or else one
   int kcap = 0;
   if (!y.reduce && kacc > 0) kcap = 2 * kacc <= budget_all / 2 ? 2 * kacc : (kacc <= budget_all / 2 ? kacc : 0);
   const int budget = budget_all - kcap;
   y.kacc = kcap ? kacc : 0;
   y.kacc_cap = kcap;
   const int fixed = (y.reduce ? need_I : 0) + CQ_PCH;
   y.ws_global = 0;
   int lb = (budget - fixed) / (y.lws + CQ_PCH);
   if (lb < 1) {
      lb = 1;
      y.ws_global = 1;
   }
   y.lb = std::min(lb, nq);
   int off = 0;
   y.o_ws = off;
   if (!y.ws_global) off += y.lb * y.lws;
   y.o_kp = off;
   off += y.lb * CQ_PCH;
   y.o_I = off;
   if (y.reduce) off += need_I;
   y.o_src = 0;
   // the Kronecker path (cq_kron): per pass of kb contractions a block of
   // bra components x ket spherical pairs each (whole 16x16 tiles), the T
   // tables, with the class's own contraction (no eco) the pass's blocks and
   // coefficient products of CQ_PCH primitive quartets, then the rows' maps
   // (in LDS, or in the workgroup's HBM slice for a class whose layered
   // workspace is there).  The layered path's LDS is free while a class
   // takes this one: its region starts at 0.  kb: at most 16 integral tiles
   // per pass (four per wavefront; the pass's contractions side by side in
   // the tile columns), at most 16, fewer where LDS is short
   int nbra = 0, nket = 0;
   for (int e = cp.la; e <= cp.LA; e++) nbra += nco(e);
   for (int f = cp.lc; f <= cp.LC; f++) nket += nco(f);
   const int nsab = (2 * cp.la + 1) * (2 * cp.lb + 1), nscd = (2 * cp.lc + 1) * (2 * cp.ld + 1);
   const int TM1 = (nbra + 15) / 16, TN = (nscd + 15) / 16, TM2 = (nsab + 15) / 16;
   const int YB = 16 * std::max(TM1, TM2) * 16 * TN;
   const int Tt = (2 * cp.ld + 1) * ioff3(cp.ld + 1) + (2 * cp.lb + 1) * ioff3(cp.lb + 1);
   const int maps = nscd * nket + nsab * nbra;
   y.o_kron = 0;
   y.kron_ws = 0;
   y.o_kmap = -1;
   y.o_kc = -1;
   y.kron = 0;
   y.kron_kb = 0;
   int kbytes = 0;
   // (kb: the pass's integral tiles, TM2 x ceil(kb nscd / 16), at most 16)
   int kb0 = std::max(1, std::min(16, nq));
   while (kb0 > 1 && TM2 * ((kb0 * nscd + 15) / 16) > 16) kb0--;
   for (int kb = kb0; kb >= 1; kb--) {
      const int klds = kb * YB + Tt;
      const int kcneed = kb * (cp.ncp + CQ_PCH);
      if (y.ws_global) {
         // the maps of up to 8 rows in the HBM slice the layered path uses
         if (klds > budget) continue;
         y.kron_ws = std::max(1, std::min(8, y.lws / maps)) * maps;
         kbytes = klds;
         if (klds + kcneed <= budget) {
            y.o_kc = klds;
            kbytes += kcneed;
         }
      } else {
         // the maps of up to 8 rows in LDS, within the type's budget
         const bool own = klds + kcneed + maps <= budget;
         const int base = klds + (own ? kcneed : 0);
         if (base + maps > budget) continue;
         y.o_kc = own ? klds : -1;
         y.o_kmap = base;
         y.kron_ws = std::min(8, (budget - base) / maps) * maps;
         kbytes = base + y.kron_ws;
      }
      y.kron_kb = kb;
      break;
   }
   y.o_kacc = kcap ? std::max(off, kbytes) : -1;
   y.bytes = 8 * (std::max(off, kbytes) + kcap);
   return y;
}

void make_cq_plan(const TypePlan &tp, const std::vector<ClassPlan> &cps, std::vector<CqLayout> &lays,
                  std::vector<int> &srcmaps, int *lds_bytes, int *gws_stride)
{
   // the budget of the type: what its largest class needs, within 50 KB
   // per workgroup all told (three workgroups per CU: more wavefronts to
   // hide the latency; a 64 KB budget, two per CU, was 8% slower on the
   // 128-water workload and 1% on the 64-water + spdf one)
   int budget = 512;
   for (int k = 0; k < tp.nclass; k++) {
      if (tq_has(tp.cls_code[k])) continue;
      const int need = 3 * cps[k].ycap + cps[k].Nc + CQ_PCH + tp.cls_nl[k] * cps[k].Ns + CQ_PCH;
      budget = std::max(budget, std::min(need, 6400 - 160 - tp.nco_tot - 4 - CQ_TAB_BYTES / 8 - CQ_PLIST_BYTES / 8 - 2));
   }
   lays.clear();
   srcmaps.clear();
   *lds_bytes = 0;
   *gws_stride = 0;
   int eco_row = 0;
   for (int k = 0; k < tp.nclass; k++) {
      const ClassPlan &cp = cps[k];
      const int sh = tp.cls_sh[k];
      const int NA = tp.nl[0][sh & 15] * (2 * cp.la + 1), NB = tp.nl[1][(sh >> 4) & 15] * (2 * cp.lb + 1),
                NC = tp.nl[2][(sh >> 8) & 15] * (2 * cp.lc + 1), ND = tp.nl[3][(sh >> 12) & 15] * (2 * cp.ld + 1);
      CqLayout y = cq_layout(cp, tp.cls_nl[k], budget, NA * NC + NA * ND + NB * NC + NB * ND);
      y.skip = tq_has(tp.cls_code[k]) ? 1 : 0;
      // many contractions: contracted on the matrix cores (k_eco), a row's
      // contraction buffer holding the class's nq x ncp doubles.  Every
      // class with at least ECO_CAND_NQ contractions has its place in the
      // buffer, so that the autotuner can move it there
      const bool cand = tp.cls_nl[k] >= ECO_CAND_NQ;
      y.eco = (cand && !y.skip && tp.cls_nl[k] >= ECO_MIN_NQ) ? 1 : 0;
      y.eco_off = cand ? eco_row : -1;
      if (cand) eco_row += tp.cls_nl[k] * cp.ncp;
      y.o_src = (int)srcmaps.size();
      for (int e = cp.la; e <= cp.LA; e++)
         for (int f = cp.lc; f <= cp.LC; f++)
            for (int i = 0; i < nco(e) * nco(f); i++) srcmaps.push_back(tp.cpoff[e][f] + i);
      // every class, so that one of the templated kernel's classes can move
      // to k_gqs
      *lds_bytes = std::max(*lds_bytes, y.bytes);
      // the Kronecker path by default where the layered one has its
      // workspace in HBM (the autotuner times both)
      y.kron = (y.eco && y.ws_global && y.kron_ws) ? 1 : 0;
      if (y.ws_global) *gws_stride = std::max(*gws_stride, std::max(y.lws, y.kron_ws));
      y.part = 0;
      lays.push_back(y);
   }
   assign_parts(tp, cps, lays);
}

// the k_gqs classes over the GQS_PARTS concurrent launches: largest
// estimated cost first, each to the lightest part
void assign_parts(const TypePlan &tp, const std::vector<ClassPlan> &cps, std::vector<CqLayout> &lays)
{
   std::vector<std::pair<double, int>> cost;
   for (int k = 0; k < tp.nclass; k++)
      if (!lays[k].skip)
         cost.push_back({(double)tp.cls_nl[k] * ((lays[k].eco ? 0.0 : 2.0 * cps[k].ncp) + cps[k].hrr_flops +
                                                  cps[k].sph_flops + 8.0 * cps[k].Ns), k});
   std::sort(cost.begin(), cost.end(), [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
      return a.first > b.first || (a.first == b.first && a.second < b.second);
   });
   double load[GQS_PARTS] = {};
   for (const auto &c : cost) {
      int best = 0;
      for (int q = 1; q < GQS_PARTS; q++)
         if (load[q] < load[best]) best = q;
      lays[c.second].part = best;
      load[best] += c.first;
   }
}

int launch_gqs(const TypePlan &tp, const ClassPlan *d_cps, const CqLayout *d_lays, const int *d_srcmaps, int lds_bytes,
               int gws_stride, int part, int npc, const DevTables &t, const BatchDev &b, const KTargets &kt, Scratch s,
               void *stream, bool wide)
{
   if (b.g1 <= b.g0) return 0;
   const int cls_bytes = (lds_bytes + 15) & ~15;
   const size_t total = (size_t)cls_bytes + CPL_BYTES + 8 * (size_t)tp.nco_tot + 4 * CQ_PCH + 16 + CQ_TAB_BYTES +
                        CQ_PLIST_BYTES;
   if (total > LDS_LIMIT) return -3;
   if (npc <= 0) return 0;
   const uint64_t nitems = (uint64_t)(b.g1 - b.g0) * (uint64_t)npc;
   if (nitems > 0xffffffffull) return -4;
   unsigned grid = (unsigned)std::min<uint64_t>(nitems, GQS_MAX_GRID);
   double *gws = nullptr;
   if (gws_stride) {
      // each part has its own workspace (the parts run concurrently), one
      // slice per workgroup
      const unsigned gcap = s.grid ? std::min(s.grid, SCRATCH_GRID) : SCRATCH_GRID;
      grid = std::min(grid, gcap);
      if (part < 0 || part >= GQS_PARTS || !s.ptr || s.doubles < (size_t)GQS_PARTS * gws_stride * gcap) return -2;
      gws = s.ptr + (size_t)part * gws_stride * gcap;
   }
   // a type whose workspace leaves room for few workgroups per CU gets wider
   // ones, so that more wavefronts hide the memory latency; so does one with
   // a workspace in HBM or a class on the Kronecker path (`wide`: that path
   // takes four wavefronts)
   // (above a third of the CU's 160 KB of LDS two workgroups fit, above a
   // quarter three: the variants with 256 and 168 VGPRs lose no wavefronts)
   if (total > 160 * 1024 / 3)
      hipLaunchKernelGGL((k_gqs<256, 2>), dim3(grid), dim3(256), total, (hipStream_t)stream, d_cps, d_lays, d_srcmaps,
                         b.tp, t, b, kt, gws, gws_stride, cls_bytes, part);
   else if (total > 160 * 1024 / 4)
      hipLaunchKernelGGL((k_gqs<256, 3>), dim3(grid), dim3(256), total, (hipStream_t)stream, d_cps, d_lays, d_srcmaps,
                         b.tp, t, b, kt, gws, gws_stride, cls_bytes, part);
   else if (total > 32 * 1024 || gws_stride > 0 || wide)
      hipLaunchKernelGGL(k_gqs<256>, dim3(grid), dim3(256), total, (hipStream_t)stream, d_cps, d_lays, d_srcmaps,
                         b.tp, t, b, kt, gws, gws_stride, cls_bytes, part);
   else
      hipLaunchKernelGGL(k_gqs<128>, dim3(grid), dim3(128), total, (hipStream_t)stream, d_cps, d_lays, d_srcmaps,
                         b.tp, t, b, kt, gws, gws_stride, cls_bytes, part);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// scratch of the no-allocation executor: one array per GPU, held for the
// life of the process.  In host code `scratch` only names the symbol; its
// device address comes from hipGetSymbolAddress, per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[NOALLOC_BYTES / sizeof(double)];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

void *noalloc_scratch()
{
   int dev = 0;
   if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
   std::lock_guard<std::mutex> lock(g_scratch_mu);
   if (!g_scratch_ptr[dev]) {
      void *ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
      g_scratch_ptr[dev] = ptr;
   }
   return g_scratch_ptr[dev];
}

__global__ void k_zero_u32(uint32_t *p, int n)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = 0u;
}

// max |P| over n elements into *out (a positive double orders as its bits;
// *out starts at 0)
__global__ void k_absmax(const double *__restrict__ P, size_t n, unsigned long long *out)
{
   double m = 0.0;
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
      m = fmax(m, fabs(P[i]));
   for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
   if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// The mixed-precision route of each set quartet of a chunk: single
// precision when its bound  max|block element| x max|coefficient|^4 x
// primitive quartets  times max|P| is below tau.  count[0] += flagged.
__global__ void k_mixed_flags(const TypePlan *__restrict__ tpp, BatchDev b, const double *__restrict__ gmag,
                              const unsigned long long *__restrict__ pmax, double tau, uint32_t *gflag,
                              unsigned long long *count)
{
   const TypePlan &tp = *tpp;
   const double maxP = __longlong_as_double((long long)*pmax);
   for (uint32_t gi = b.g0 + blockIdx.x * blockDim.x + threadIdx.x; gi < b.g1; gi += gridDim.x * blockDim.x) {
      const GroupRec &g = b.groups[gi];
      double c4 = 1.0;
      for (int t = 0; t < 4; t++) {
         const int np = ((g.np4 >> (4 * t)) & 15) + 1;
         double cm = 0.0;
         for (int sh = 0; sh < tp.nsh[t]; sh++)
            for (int i = 0; i < tp.nl[t][sh] * np; i++) cm = fmax(cm, fabs(b.data[g.Kof[t][sh] + i]));
         c4 *= cm;
      }
      double npr = 0.0;
      for (uint32_t r = g.row0; r < g.row1; r++) npr += (double)b.rows[r].n_prm;
      const bool f32 = gmag[gi] * c4 * npr * maxP < tau;
      gflag[gi] = f32 ? 1u : 0u;
      if (f32) atomicAdd(count, 1ull);
   }
}

int launch_absmax(const double *P, size_t n, unsigned long long *out, bool zero, void *stream)
{
   if (zero) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, (hipStream_t)stream, (uint32_t *)out, 2);
   const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 2048);
   hipLaunchKernelGGL(k_absmax, dim3(std::max(grid, 1u)), dim3(256), 0, (hipStream_t)stream, P, n, out);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mixed_flags(const TypePlan &tp, const BatchDev &b, const double *gmag, const unsigned long long *pmax,
                       double tau, uint32_t *gflag, unsigned long long *count, void *stream)
{
   (void)tp;
   if (b.g1 <= b.g0) return 0;
   const unsigned grid = std::min<unsigned>((b.g1 - b.g0 + 255) / 256, 4096);
   hipLaunchKernelGGL(k_mixed_flags, dim3(grid), dim3(256), 0, (hipStream_t)stream, b.tp, b, gmag, pmax, tau, gflag,
                      count);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_zero_f64(double *p, size_t n)
{
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

void launch_zero_f64(double *p, size_t n, void *stream)
{
   if (!n) return;
   const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
   hipLaunchKernelGGL(k_zero_f64, dim3(grid), dim3(256), 0, (hipStream_t)stream, p, n);
}

int launch_zero_u32(uint32_t *p, int n, void *stream)
{
   if (n <= 0) return 0;
   hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, n);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// self-test of the screening arithmetic (bitwise comparison with the host)
// ---------------------------------------------------------------------------
__global__ void k_math_selftest(const double *x, const double *y, int n, double *out)
{
   GINT_NO_CONTRACT
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const double a = x[i], c = y[i];
   out[4 * i + 0] = gint_exp(a);
   out[4 * i + 1] = sqrt(fabs(a));
   out[4 * i + 2] = 1. / a;
   out[4 * i + 3] = a * c / (a + c);
}

int device_math_selftest(const double *x, const double *y, int n, double *out)
{
   double *dx = nullptr, *dy = nullptr, *dout = nullptr;
   int rc = 0;
   if (hipMalloc(&dx, sizeof(double) * n) != hipSuccess || hipMalloc(&dy, sizeof(double) * n) != hipSuccess ||
       hipMalloc(&dout, sizeof(double) * 4 * n) != hipSuccess)
      rc = -1;
   if (!rc && (hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dy, y, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess))
      rc = -1;
   if (!rc) {
      hipLaunchKernelGGL(k_math_selftest, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dy, n, dout);
      if (hipGetLastError() != hipSuccess) rc = -1;
   }
   if (!rc && hipMemcpy(out, dout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
   if (dx) (void)hipFree(dx);
   if (dy) (void)hipFree(dy);
   if (dout) (void)hipFree(dout);
   return rc;
}

}  // namespace gint
