// gfx950 kernels of the K-build, one angular-momentum class per launch.
//
//   k_fm_vrr       one wavefront per primitive quartet.  Lanes evaluate
//                  G_m(R,T) (or F_m) for 64 lattice cells at once, then the
//                  wavefront runs the Obara-Saika VRR of every surviving cell
//                  in LDS and sums the (e0|f0) blocks over cells on chip.
//                  Replaces reference src/compute_Fm.cpp (prepare, Fm, Vm),
//                  src/compute_VRR.cpp and the cell sum of
//                  src/compute_ECO.cpp:216-272; nothing per cell touches HBM.
//   k_eco          one workgroup per row: contraction over primitives,
//                  reference src/compute_ECO.cpp:12-214.
//   k_hrr_sph_ks   one wavefront per contracted quartet: HRR of each (n1,n2)
//                  row, summed on chip (no atomics), cartesian -> spherical,
//                  exchange update of K.  Reference src/compute_HRR.cpp,
//                  src/compute_SPH.cpp:58-237, src/compute_KS.cpp:8-117.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "gint_math.h"
#include "c2s_table.h"

namespace gint {

// ---------------------------------------------------------------------------
// cartesian component tables for shells 0..7, built at compile time
// ---------------------------------------------------------------------------
struct CartTables {
   unsigned char comp[8][36][3];   // lx ly lz of component i of shell L
   unsigned char first[8][36];     // first non-zero direction
   signed char minus[8][36][3];    // index of i - 1_d in shell L-1, -1 if none
   signed char plus[8][36][3];     // index of i + 1_d in shell L+1
};

constexpr int cart_index(int L, int lx, int lz) { return (L - lx) * (L - lx + 1) / 2 + lz; }

constexpr CartTables make_cart_tables()
{
   CartTables t{};
   for (int L = 0; L < 8; L++) {
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
               int u[3] = {v[0], v[1], v[2]};
               u[d]++;
               t.plus[L][i][d] = (signed char)cart_index(L + 1, u[0], u[2]);
            }
            i++;
         }
   }
   return t;
}

__constant__ CartTables c_cart = make_cart_tables();
__constant__ double c_c2s[GINT_C2S_SIZE] = {GINT_C2S_VALUES};
__constant__ int c_c2s_off[4] = {GINT_C2S_OFF_0, GINT_C2S_OFF_1, GINT_C2S_OFF_2, GINT_C2S_OFF_3};

__device__ __forceinline__ int d_nco(int l) { return (l + 1) * (l + 2) / 2; }

// ---------------------------------------------------------------------------
// per-primitive quantities, in the order of reference
// src/compute_Fm.cpp:66-159 and 202-210, 281-290 (no contraction: these
// feed the Fm[0] > 1e-12 decision)
// ---------------------------------------------------------------------------
struct PrimGeo {
   double P[3], Q[3], PA[3], QC[3];
   double Kfac, inv_z, zab, zcd, rho, Rr;
   double i2ab, mrab, i2cd, mrcd, i2z;
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
   double A[3], B[3], C[3], D[3], ABs[3], CDs[3], PQs[3], shift[3], Cs[3], Ds[3], P[3], Q[3];
   gint_pbc(Ao, Bo, t.cell, ABs);
   A[0] = Ao[0]; A[1] = Ao[1]; A[2] = Ao[2];
   B[0] = Ao[0] + ABs[0] + neighs[n1 * 3 + 0];
   B[1] = Ao[1] + ABs[1] + neighs[n1 * 3 + 1];
   B[2] = Ao[2] + ABs[2] + neighs[n1 * 3 + 2];
   P[0] = (za * A[0] + zb * B[0]) * inv_zab;
   P[1] = (za * A[1] + zb * B[1]) * inv_zab;
   P[2] = (za * A[2] + zb * B[2]) * inv_zab;
   gint_pbc(Co, Do, t.cell, CDs);
   C[0] = Co[0]; C[1] = Co[1]; C[2] = Co[2];
   D[0] = Co[0] + CDs[0] + neighs[n2 * 3 + 0];
   D[1] = Co[1] + CDs[1] + neighs[n2 * 3 + 1];
   D[2] = Co[2] + CDs[2] + neighs[n2 * 3 + 2];
   Q[0] = (zc * C[0] + zd * D[0]) * inv_zcd;
   Q[1] = (zc * C[1] + zd * D[1]) * inv_zcd;
   Q[2] = (zc * C[2] + zd * D[2]) * inv_zcd;
   gint_pbc(Q, P, t.cell, PQs);
   shift[0] = P[0] - Q[0] - PQs[0];
   shift[1] = P[1] - Q[1] - PQs[1];
   shift[2] = P[2] - Q[2] - PQs[2];
   const double normf = GINT_TWO_PI_52 * inv_zab * inv_zcd;
   double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
   const double RAB2 = dx * dx + dy * dy + dz * dz;
   const double expon_ab = -(za * zb) / (za + zb) * RAB2;
   dx = C[0] - D[0]; dy = C[1] - D[1]; dz = C[2] - D[2];
   const double RCD2 = dx * dx + dy * dy + dz * dz;
   const double expon_cd = -(zc * zd) / (zc + zd) * RCD2;
   const double Zn = 1. / sqrt(z);
   g.Kfac = Zn * normf * gint_exp(expon_ab + expon_cd);
   Cs[0] = Co[0] + shift[0];
   Cs[1] = Co[1] + shift[1];
   Cs[2] = Co[2] + shift[2];
   Ds[0] = Co[0] + CDs[0] + neighs[n2 * 3 + 0] + shift[0];
   Ds[1] = Co[1] + CDs[1] + neighs[n2 * 3 + 1] + shift[1];
   Ds[2] = Co[2] + CDs[2] + neighs[n2 * 3 + 2] + shift[2];
   Q[0] = (zc * Cs[0] + zd * Ds[0]) * inv_zcd;
   Q[1] = (zc * Cs[1] + zd * Ds[1]) * inv_zcd;
   Q[2] = (zc * Cs[2] + zd * Ds[2]) * inv_zcd;
   for (int k = 0; k < 3; k++) {
      g.P[k] = P[k];
      g.Q[k] = Q[k];
      g.PA[k] = P[k] - A[k];
      g.QC[k] = Q[k] - Cs[k];
   }
   g.inv_z = inv_z;
   g.zab = zab;
   g.zcd = zc + zd;
   g.rho = zab * g.zcd * inv_z;
   g.Rr = t.R_cut * sqrt(g.rho);
   const double izab = 1. / zab;
   g.i2ab = 0.5 * izab;
   g.mrab = -1. * g.rho * izab * g.i2ab;
   const double izcd = 1. / g.zcd;
   g.i2cd = 0.5 * izcd;
   g.mrcd = -1. * g.rho * izcd * g.i2cd;
   g.i2z = 0.5 * inv_z;
}

// Fm[0..L] of one cell, reference src/compute_Fm.cpp:212-236 and
// src/t_c_g0_n.cpp:67-180.  Fm[0] is evaluated without contraction.
__device__ void slot_fm(const DevTables &t, int L, const PrimGeo &g, int n3, double *Fm)
{
   double T;
   {
      GINT_NO_CONTRACT
      double PQ0 = g.P[0] - g.Q[0] - t.neighs[n3 * 3 + 0];
      double PQ1 = g.P[1] - g.Q[1] - t.neighs[n3 * 3 + 1];
      double PQ2 = g.P[2] - g.Q[2] - t.neighs[n3 * 3 + 2];
      double rpq2 = PQ0 * PQ0 + PQ1 * PQ1 + PQ2 * PQ2;
      T = g.rho * rpq2;
   }
   if (!t.truncated) {
      gint_boys(L, T, Fm, t.ftable, g.Kfac);
      return;
   }
   int patch;
   double TG1, TG2;
   int br = gint_tcg_locate(g.Rr, T, t.grid_lo, t.grid_hi, t.bw, &patch, &TG1, &TG2);
   if (br == GINT_G_ZERO) {
      for (int k = 0; k <= L; k++) Fm[k] = 0.0;
      return;
   }
   if (br == GINT_G_BOYS) {
      gint_boys(L, T, Fm, t.ftable, g.Kfac);
      return;
   }
   double T1[14], T2[14];
   gint_cheb14(TG1, T1);
   gint_cheb14(TG2, T2);
   const double *row = t.C0 + (size_t)t.ld_C0 * patch;
   {
      GINT_NO_CONTRACT
      Fm[0] = gint_cheb_sum(row, T1, T2) * g.Kfac;
   }
   for (int k = 1; k <= L; k++) {
      const double *c = row + k * TCG_NCOEF;
      double tmp = 0.0;
      int jl = 0;
      for (int j = 0; j < 14; j++) {
         double dot = 0.0;
         for (int l = 0; l < 14 - j; l++) dot += T1[l] * c[jl++];
         tmp += dot * T2[j];
      }
      Fm[k] = tmp * g.Kfac;
   }
}

// order-independent survivor checksum, same function as the oracle's
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// VRR of one slot in the workspace, lanes over elements.
// ws block (a,c) at voff[a][c], layout [m][ia][ic], m = 0..L-a-c.
// pq: PA WP QC WQ (3 each) i2ab mrab i2cd mrcd i2z
// ---------------------------------------------------------------------------
__device__ void vrr_slot(const ClassPlan &cp, double *ws, const double *Fm, const double *pq, int lane, int nlane)
{
   const int L = cp.L;
   for (int m = lane; m <= L; m += nlane) ws[cp.voff[0][0] + m] = Fm[m];
   __syncthreads();
   const double i2ab = pq[12], mrab = pq[13], i2cd = pq[14], mrcd = pq[15], i2z = pq[16];
   for (int a = 1; a <= cp.LA; a++) {
      const int na = d_nco(a), na1 = d_nco(a - 1), na2 = a >= 2 ? d_nco(a - 2) : 1;
      const int n = (L - a + 1) * na;
      const double *V1 = ws + cp.voff[a - 1][0];
      const double *V2 = ws + (a >= 2 ? cp.voff[a - 2][0] : 0);
      double *V = ws + cp.voff[a][0];
      for (int idx = lane; idx < n; idx += nlane) {
         const int m = idx / na, ia = idx - m * na;
         const int d = c_cart.first[a][ia];
         const int i1 = c_cart.minus[a][ia][d];
         const int e2 = c_cart.comp[a][ia][d] - 1;
         double val = pq[d] * V1[m * na1 + i1] + pq[3 + d] * V1[(m + 1) * na1 + i1];
         if (e2 > 0) {
            const int i2 = c_cart.minus[a - 1][i1][d];
            val += e2 * (i2ab * V2[m * na2 + i2] + mrab * V2[(m + 1) * na2 + i2]);
         }
         V[idx] = val;
      }
      __syncthreads();
   }
   for (int c = 1; c <= cp.LC; c++) {
      const int nc = d_nco(c), nc1 = d_nco(c - 1), nc2 = c >= 2 ? d_nco(c - 2) : 1;
      int total = 0;
      for (int a = cp.amin[c]; a <= cp.LA; a++) total += (L - a - c + 1) * d_nco(a) * nc;
      for (int idx0 = lane; idx0 < total; idx0 += nlane) {
         int idx = idx0, a = cp.amin[c];
         for (;; a++) {
            const int sz = (L - a - c + 1) * d_nco(a) * nc;
            if (idx < sz) break;
            idx -= sz;
         }
         const int na = d_nco(a);
         const int m = idx / (na * nc);
         const int rem = idx - m * na * nc;
         const int ia = rem / nc, ic = rem - ia * nc;
         const int d = c_cart.first[c][ic];
         const int j1 = c_cart.minus[c][ic][d];
         const int f2 = c_cart.comp[c][ic][d] - 1;
         const double *Vc1 = ws + cp.voff[a][c - 1];
         double val = pq[6 + d] * Vc1[(m * na + ia) * nc1 + j1] + pq[9 + d] * Vc1[((m + 1) * na + ia) * nc1 + j1];
         if (f2 > 0) {
            const int j2 = c_cart.minus[c - 1][j1][d];
            const double *Vc2 = ws + cp.voff[a][c - 2];
            val += f2 * (i2cd * Vc2[(m * na + ia) * nc2 + j2] + mrcd * Vc2[((m + 1) * na + ia) * nc2 + j2]);
         }
         const int e2 = c_cart.comp[a][ia][d];
         if (e2 > 0) {
            const int i1 = c_cart.minus[a][ia][d];
            const int na1 = d_nco(a - 1);
            const double *Va = ws + cp.voff[a - 1][c - 1];
            val += e2 * i2z * Va[((m + 1) * na1 + i1) * nc1 + j1];
         }
         ws[cp.voff[a][c] + (m * na + ia) * nc + ic] = val;
      }
      __syncthreads();
   }
}

// ---------------------------------------------------------------------------
// K1: Fm -> VRR -> sum over cells, one 64-lane workgroup per primitive
// LDS: [ws vws (if in LDS)] [fm 64 x (L+1)] [acc ncp]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_fm_vrr(ClassPlan cp, DevTables t, BatchDev b, double *gws)
{
   extern __shared__ double lds[];
   const int lane = threadIdx.x;
   const int L = cp.L;
   double *ws = gws ? gws + (size_t)blockIdx.x * cp.vws : lds;
   double *fmb = gws ? lds : lds + cp.vws;
   // cell sums: acc0 takes cell 0, acc the others; out = acc0 + acc, the
   // order of reference src/compute_ECO.cpp:263-269
   double *acc = fmb + 64 * (L + 1);
   double *acc0 = acc + cp.ncp;
   const bool ssss_src = (cp.la == 0 && cp.lc == 0);
   unsigned long long n_surv = 0, csum = 0;
   const unsigned long long wcls = (unsigned long long)((cp.la << 12) | (cp.lb << 8) | (cp.lc << 4) | cp.ld);

   for (uint32_t p = blockIdx.x; p < b.n_prm; p += gridDim.x) {
      const uint32_t r = b.prm_row[p];
      const RowRec row = b.rows[r];
      const uint32_t ip = b.prm_ip[p];
      const int ipa = ip & 15, ipb = (ip >> 4) & 15, ipc = (ip >> 8) & 15, ipd = (ip >> 12) & 15;
      const double *data = b.data;
      const int n1 = row.n12 & 0xffff, n2 = row.n12 >> 16;
      PrimGeo g;
      prim_prepare(data + row.R[0], data + row.R[1], data + row.R[2], data + row.R[3],
                   data[row.Z[0] + ipa], data[row.Z[1] + ipb], data[row.Z[2] + ipc], data[row.Z[3] + ipd],
                   n1, n2, t, g);
      for (int j = lane; j < cp.ncp; j += 64) acc[j] = acc0[j] = 0.0;

      for (int base = 0; base < t.Ng; base += 64) {
         const int n3 = base + lane;
         bool surv = false;
         if (n3 < t.Ng) {
            double Fm[MAXLL + 1];
            slot_fm(t, L, g, n3, Fm);
            for (int k = 0; k <= L; k++) fmb[lane * (L + 1) + k] = Fm[k];
            surv = Fm[0] > 1.e-12;
         }
         const unsigned long long mask = __ballot(surv);
         if (surv && b.stats) {
            // key of the slot, same mixing as the oracle's slot key
            const unsigned long long w0 = ((unsigned long long)(row.set_ij & 0xffff) << 32) | (row.set_ij >> 16);
            const unsigned long long w1 = ((unsigned long long)(row.set_kl & 0xffff) << 32) | (row.set_kl >> 16);
            const unsigned long long w2 = ((unsigned long long)n1 << 24) | ((unsigned long long)n2 << 16) | wcls;
            const unsigned long long w3 = ((unsigned long long)((ipa << 12) | (ipb << 8) | (ipc << 4) | ipd) << 8) |
                                          (unsigned long long)n3;
            unsigned long long z = mix64(w0 + 0x9E3779B97F4A7C15ULL);
            z = mix64(z ^ w1);
            z = mix64(z ^ w2);
            z = mix64(z ^ w3);
            n_surv++;
            csum += z;
         }
         __syncthreads();
         const int nslot = min(64, t.Ng - base);
         for (int s = 0; s < nslot; s++) {
            const double *Fm = fmb + s * (L + 1);
            if ((mask >> s) & 1ull) {
               const int n3s = base + s;
               double pq[17];
               {
                  const double gx = t.neighs[n3s * 3 + 0], gy = t.neighs[n3s * 3 + 1], gz = t.neighs[n3s * 3 + 2];
                  const double Qp[3] = {g.Q[0] + gx, g.Q[1] + gy, g.Q[2] + gz};
                  double W[3];
                  for (int k = 0; k < 3; k++) W[k] = (g.zab * g.P[k] + g.zcd * Qp[k]) * g.inv_z;
                  for (int k = 0; k < 3; k++) {
                     pq[k] = g.PA[k];
                     pq[3 + k] = W[k] - g.P[k];
                     pq[6 + k] = g.QC[k];
                     pq[9 + k] = W[k] - Qp[k];
                  }
                  pq[12] = g.i2ab; pq[13] = g.mrab; pq[14] = g.i2cd; pq[15] = g.mrcd; pq[16] = g.i2z;
               }
               vrr_slot(cp, ws, Fm, pq, lane, 64);
               double *a = (n3s == 0) ? acc0 : acc;
               for (int e = cp.la; e <= cp.LA; e++)
                  for (int f = cp.lc; f <= cp.LC; f++) {
                     const int sz = d_nco(e) * d_nco(f);
                     const double *src = ws + cp.voff[e][f];
                     double *dst = a + cp.cpoff[e][f];
                     for (int k = lane; k < sz; k += 64) dst[k] += src[k];
                  }
               __syncthreads();
            } else if (ssss_src && lane == 0 && base + s < t.Ng) {
               // a screened slot still carries its ssss(0) value
               // (reference src/compute_VRR.cpp:384-385)
               if (base + s == 0) acc0[0] += Fm[0];
               else acc[0] += Fm[0];
            }
         }
         __syncthreads();
      }
      double *out = b.ac0 + (size_t)p * cp.ncp;
      for (int j = lane; j < cp.ncp; j += 64) out[j] = acc0[j] + acc[j];
      __syncthreads();
   }
   if (b.stats) {
      for (int off = 32; off > 0; off >>= 1) {
         n_surv += __shfl_down(n_surv, off, 64);
         csum += __shfl_down(csum, off, 64);
      }
      if (lane == 0 && n_surv) {
         atomicAdd(&b.stats[0], n_surv);
         atomicAdd(&b.stats[1], csum);
         atomicAdd(&b.stats[2], n_surv * (unsigned long long)(cp.fm_flops + cp.vrr_flops));
      }
   }
}

// ---------------------------------------------------------------------------
// K2: contraction over primitives, one 256-thread workgroup per row.
// out[(e_off + l) * ncp + j] = sum_p Ka Kb Kc Kd (p) * ac0[p][j]
// ---------------------------------------------------------------------------
constexpr int ECO_THREADS = 256;
constexpr int ECO_PCHUNK = 16;
constexpr int ECO_PER_THREAD = 4;

__global__ __launch_bounds__(ECO_THREADS) void k_eco(ClassPlan cp, BatchDev b, uint32_t row_begin, uint32_t row_end,
                                                      uint32_t prm_begin, uint32_t prm_end)
{
   __shared__ double sK[MAXNL * MAXNL * MAXNL * MAXNL * ECO_PCHUNK];
   const int tid = threadIdx.x;
   const int ncp = cp.ncp;
   for (uint32_t r = row_begin + blockIdx.x; r < row_end; r += gridDim.x) {
      const RowRec row = b.rows[r];
      const uint32_t pa = row.prm_off > prm_begin ? row.prm_off : prm_begin;
      const uint32_t pb = row.prm_off + row.n_prm < prm_end ? row.prm_off + row.n_prm : prm_end;
      if (pa >= pb) continue;
      const int nla = row.nlnp & 15, nlb = (row.nlnp >> 4) & 15, nlc = (row.nlnp >> 8) & 15, nld = (row.nlnp >> 12) & 15;
      const int npa = ((row.nlnp >> 16) & 15) + 1, npb = ((row.nlnp >> 20) & 15) + 1;
      const int npc = ((row.nlnp >> 24) & 15) + 1, npd = ((row.nlnp >> 28) & 15) + 1;
      const int nlabcd = nla * nlb * nlc * nld;
      const double *Ka = b.data + row.K[0], *Kb = b.data + row.K[1], *Kc = b.data + row.K[2], *Kd = b.data + row.K[3];
      const int total = nlabcd * ncp;
      double *out = b.eco + (size_t)row.e_off * ncp;
      for (int tile = 0; tile < total; tile += ECO_THREADS * ECO_PER_THREAD) {
         double accv[ECO_PER_THREAD] = {0.0, 0.0, 0.0, 0.0};
         for (uint32_t p0 = pa; p0 < pb; p0 += ECO_PCHUNK) {
            const int np = (int)min((uint32_t)ECO_PCHUNK, pb - p0);
            __syncthreads();
            for (int idx = tid; idx < nlabcd * ECO_PCHUNK; idx += ECO_THREADS) {
               const int l = idx / ECO_PCHUNK, pp = idx - l * ECO_PCHUNK;
               double kv = 0.0;
               if (pp < np) {
                  const uint32_t ip = b.prm_ip[p0 + pp];
                  const int ipa = ip & 15, ipb = (ip >> 4) & 15, ipc = (ip >> 8) & 15, ipd = (ip >> 12) & 15;
                  const int a = l / (nlb * nlc * nld), bb = (l / (nlc * nld)) % nlb, c = (l / nld) % nlc, d = l % nld;
                  kv = Ka[a * npa + ipa] * Kb[bb * npb + ipb] * Kc[c * npc + ipc] * Kd[d * npd + ipd];
               }
               sK[idx] = kv;
            }
            __syncthreads();
            const double *A = b.ac0 + (size_t)(p0 - prm_begin) * ncp;
            for (int u = 0; u < ECO_PER_THREAD; u++) {
               const int idx = tile + u * ECO_THREADS + tid;
               if (idx < total) {
                  const int l = idx / ncp, j = idx - l * ncp;
                  double s = accv[u];
                  for (int pp = 0; pp < np; pp++) s += sK[l * ECO_PCHUNK + pp] * A[(size_t)pp * ncp + j];
                  accv[u] = s;
               }
            }
         }
         for (int u = 0; u < ECO_PER_THREAD; u++) {
            const int idx = tile + u * ECO_THREADS + tid;
            if (idx < total) out[idx] += accv[u];
         }
      }
   }
}

// ---------------------------------------------------------------------------
// K3: HRR over the rows of a quartet, sum, cart->sph, K update.
// workspace: three buffers of cap doubles + acc (Nc)
// ---------------------------------------------------------------------------
constexpr int Q_THREADS = 64;

__device__ void row_ab_cd(const RowRec &row, const double *data, const DevTables &t, double AB[3], double CD[3])
{
   const double *Ao = data + row.R[0], *Bo = data + row.R[1], *Co = data + row.R[2], *Do = data + row.R[3];
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

__global__ __launch_bounds__(Q_THREADS) void k_hrr_sph_ks(ClassPlan cp, DevTables t, BatchDev b, KTargets kt, double *gws)
{
   extern __shared__ double lds[];
   const int tid = threadIdx.x;
   const int cap = cp.ycap;
   double *ws = gws ? gws + (size_t)blockIdx.x * cp.hrr_ws : lds;
   double *buf[3] = {ws, ws + cap, ws + 2 * cap};
   double *acc = ws + 3 * cap;
   const int la = cp.la, lb = cp.lb, lc = cp.lc, ld = cp.ld, LA = cp.LA, LC = cp.LC;
   const int Nc = cp.Nc;
   const int ncd = d_nco(lc) * d_nco(ld);

   for (uint32_t q = blockIdx.x; q < b.n_q; q += gridDim.x) {
      const KsRec ks = b.ks[q];
      for (int i = tid; i < Nc; i += Q_THREADS) acc[i] = 0.0;
      for (uint32_t r = ks.row0; r < ks.row1; r++) {
         const RowRec row = b.rows[r];
         double AB[3], CD[3];
         row_ab_cd(row, b.data, t, AB, CD);
         const double *E = b.eco + (size_t)(row.e_off + ks.l) * cp.ncp;
         __syncthreads();
         for (int i = tid; i < cp.ncp; i += Q_THREADS) buf[0][i] = E[i];
         __syncthreads();
         int src = 0;
         // ket HRR (d -> c): layer dd from layer dd-1, blocks (e,c) e outer
         for (int dd = 1; dd <= ld; dd++) {
            const int dst = (src + 1) % 3;
            const int nd = d_nco(dd), ndm = d_nco(dd - 1);
            for (int idx0 = tid; idx0 < cp.ksize[dd]; idx0 += Q_THREADS) {
               int e = la, c = lc;
               // locate the block (blocks are few: at most 4 x 4)
               for (;;) {
                  const int end_ec = cp.koff[dd][e][c] + d_nco(e) * d_nco(c) * nd;
                  if (idx0 < end_ec) break;
                  if (++c > LC - dd) { c = lc; e++; }
               }
               const int idx = idx0 - cp.koff[dd][e][c];
               const int ncc = d_nco(c), ncc1 = d_nco(c + 1);
               const int ie = idx / (ncc * nd);
               const int rem = idx - ie * ncc * nd;
               const int ic = rem / nd, id = rem - ic * nd;
               const int dir = c_cart.first[dd][id];
               const int idm = c_cart.minus[dd][id][dir];
               const int icp = c_cart.plus[c][ic][dir];
               const double v1 = buf[src][cp.koff[dd - 1][e][c + 1] + (ie * ncc1 + icp) * ndm + idm];
               const double v0 = buf[src][cp.koff[dd - 1][e][c] + (ie * ncc + ic) * ndm + idm];
               buf[dst][idx0] = v1 + CD[dir] * v0;
            }
            __syncthreads();
            src = dst;
         }
         // bra HRR (b -> a): layer bb from layer bb-1, blocks a in [la, LA-bb]
         for (int bb = 1; bb <= lb; bb++) {
            const int dst = (src + 1) % 3;
            const int nb = d_nco(bb), nbm = d_nco(bb - 1);
            for (int idx0 = tid; idx0 < cp.bsize[bb]; idx0 += Q_THREADS) {
               int a = la;
               while (idx0 >= cp.boff[bb][a] + d_nco(a) * nb * ncd) a++;
               const int idx = idx0 - cp.boff[bb][a];
               const int ia = idx / (nb * ncd);
               const int rem = idx - ia * nb * ncd;
               const int ib = rem / ncd, icd = rem - ib * ncd;
               const int dir = c_cart.first[bb][ib];
               const int ibm = c_cart.minus[bb][ib][dir];
               const int iap = c_cart.plus[a][ia][dir];
               const double v1 = buf[src][cp.boff[bb - 1][a + 1] + (iap * nbm + ibm) * ncd + icd];
               const double v0 = buf[src][cp.boff[bb - 1][a] + (ia * nbm + ibm) * ncd + icd];
               buf[dst][idx0] = v1 + AB[dir] * v0;
            }
            __syncthreads();
            src = dst;
         }
         // the first block of the last layer is (la lb|lc ld)
         for (int i = tid; i < Nc; i += Q_THREADS) acc[i] += buf[src][i];
      }
      __syncthreads();
      // cartesian -> spherical: four stages, each moves the leading index last
      {
         const int ls[4] = {la, lb, lc, ld};
         int dims[4] = {d_nco(la), d_nco(lb), d_nco(lc), d_nco(ld)};
         const double *in = acc;
         for (int s = 0; s < 4; s++) {
            double *o = buf[s & 1];
            const int k = dims[0];
            const int n = dims[1] * dims[2] * dims[3];
            const int m = 2 * ls[s] + 1;
            const double *c2s = c_c2s + c_c2s_off[ls[s]];
            for (int idx = tid; idx < n * m; idx += Q_THREADS) {
               const int j = idx / m, i = idx - j * m;
               double v = 0.0;
               for (int qq = 0; qq < k; qq++) v += c2s[i * k + qq] * in[qq * n + j];
               o[idx] = v;
            }
            __syncthreads();
            in = o;
            const int nd0 = dims[1], nd1 = dims[2], nd2 = dims[3];
            dims[0] = nd0; dims[1] = nd1; dims[2] = nd2; dims[3] = m;
         }
      }
      const double *I = buf[1];
      // exchange update, reference src/compute_KS.cpp:51-115
      {
         const int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1, nsd = 2 * ld + 1;
         const int nscd = nsc * nsd, nsbcd = nsb * nscd, nsabcd = nsa * nsbcd;
         const double symm[5] = {2.0, 1.0, 0.5, 0.25, 0.0};
         const double fac = symm[(ks.flags >> 4) & 7] * kt.hf_fac;
         const unsigned ldac = ks.ld01 & 0xffff, ldad = ks.ld01 >> 16, ldbc = ks.ld23 & 0xffff, ldbd = ks.ld23 >> 16;
         for (int tt = tid; tt < nsabcd; tt += Q_THREADS) {
            const unsigned isa = tt / nsbcd, isb = (tt / nscd) % nsb, isc = (tt / nsd) % nsc, isd = tt % nsd;
            const unsigned iac = !(ks.flags & 1) ? isc * ldac + isa : isa * ldac + isc;
            const unsigned iad = !(ks.flags & 2) ? isd * ldad + isa : isa * ldad + isd;
            const unsigned ibc = !(ks.flags & 4) ? isc * ldbc + isb : isb * ldbc + isc;
            const unsigned ibd = !(ks.flags & 8) ? isd * ldbd + isb : isb * ldbd + isd;
            const size_t Iac = (size_t)ks.off[0] + iac, Iad = (size_t)ks.off[1] + iad;
            const size_t Ibc = (size_t)ks.off[2] + ibc, Ibd = (size_t)ks.off[3] + ibd;
            const double iabcd = -fac * I[tt];
            for (int s = 0; s < kt.nspin; s++) {
               const double *P = kt.P[s];
               double *K = kt.K[s];
               unsafeAtomicAdd(&K[Ibd], iabcd * P[Iac]);
               unsafeAtomicAdd(&K[Ibc], iabcd * P[Iad]);
               unsafeAtomicAdd(&K[Iad], iabcd * P[Ibc]);
               unsafeAtomicAdd(&K[Iac], iabcd * P[Ibd]);
            }
         }
      }
      __syncthreads();
   }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static constexpr size_t LDS_LIMIT = 64 * 1024;

static size_t vrr_lds_full(const ClassPlan &cp) { return sizeof(double) * ((size_t)cp.vws + 64 * (cp.L + 1) + 2 * cp.ncp); }
static size_t vrr_lds_part(const ClassPlan &cp) { return sizeof(double) * ((size_t)64 * (cp.L + 1) + 2 * cp.ncp); }
static size_t hrr_lds_full(const ClassPlan &cp) { return sizeof(double) * (size_t)cp.hrr_ws; }

size_t scratch_doubles(const ClassPlan &cp)
{
   size_t need = 0;
   if (vrr_lds_full(cp) > LDS_LIMIT) need = std::max(need, (size_t)cp.vws * SCRATCH_GRID);
   if (hrr_lds_full(cp) > LDS_LIMIT) need = std::max(need, (size_t)cp.hrr_ws * SCRATCH_GRID);
   return need;
}

int launch_fm_vrr(const ClassPlan &cp, const DevTables &t, const BatchDev &b, Scratch s, void *stream)
{
   if (b.n_prm == 0) return 0;
   hipStream_t st = (hipStream_t)stream;
   if (vrr_lds_full(cp) <= LDS_LIMIT) {
      hipLaunchKernelGGL(k_fm_vrr, dim3(b.n_prm), dim3(64), vrr_lds_full(cp), st, cp, t, b, (double *)nullptr);
   } else {
      const unsigned grid = b.n_prm < SCRATCH_GRID ? b.n_prm : SCRATCH_GRID;
      if (!s.ptr || s.doubles < (size_t)cp.vws * grid) return -2;
      hipLaunchKernelGGL(k_fm_vrr, dim3(grid), dim3(64), vrr_lds_part(cp), st, cp, t, b, s.ptr);
   }
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_eco_range(const ClassPlan &cp, const BatchDev &b, uint32_t row_begin, uint32_t row_end,
                     uint32_t prm_begin, uint32_t prm_end, void *stream)
{
   if (row_end <= row_begin || prm_end <= prm_begin) return 0;
   hipLaunchKernelGGL(k_eco, dim3(row_end - row_begin), dim3(ECO_THREADS), 0, (hipStream_t)stream, cp, b, row_begin,
                      row_end, prm_begin, prm_end);
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

int launch_hrr_sph_ks(const ClassPlan &cp, const DevTables &t, const BatchDev &b, const KTargets &kt, Scratch s,
                      void *stream)
{
   if (b.n_q == 0) return 0;
   hipStream_t st = (hipStream_t)stream;
   if (hrr_lds_full(cp) <= LDS_LIMIT) {
      hipLaunchKernelGGL(k_hrr_sph_ks, dim3(b.n_q), dim3(Q_THREADS), hrr_lds_full(cp), st, cp, t, b, kt,
                         (double *)nullptr);
   } else {
      const unsigned grid = b.n_q < SCRATCH_GRID ? b.n_q : SCRATCH_GRID;
      if (!s.ptr || s.doubles < (size_t)cp.hrr_ws * grid) return -2;
      hipLaunchKernelGGL(k_hrr_sph_ks, dim3(grid), dim3(Q_THREADS), 0, st, cp, t, b, kt, s.ptr);
   }
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace gint
