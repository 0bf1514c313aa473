// Templated quartet kernels for the classes whose (e0|f0) blocks have at most
// 128 columns, one instantiation per class (la lb|lc ld).  They contract each (n1,n2) row's
// primitive blocks (reference src/compute_ECO.cpp:12-214), run the HRR of the
// row (d -> c, then b -> a, src/compute_HRR.cpp), sum the rows, transform to
// spherical (src/compute_SPH.cpp:58-237) and update K (src/compute_KS.cpp:
// 51-115) with compile-time indices and c2s coefficients, so zero
// coefficients cost nothing.
//
// k_tqw: one wavefront per set quartet.  The row's primitive quartets are
// walked in step on all lanes: lane (c, g) loads column c of each block once
// (coalesced; columns c and c + 64 above 64 columns) and accumulates it into
// the contractions g, g + LL, g + 2 LL,
// ... with the coefficient products of the batch, formed once per batch in
// LDS.  The contracted blocks then go to LDS, one row per contraction (odd
// compile-time strides, so element offsets are immediates), where each lane
// runs the HRR, the row sum and the spherical transform of its own
// contractions.  The K updates of the set quartet are summed over its
// contractions before the atomics (one atomic per K element of the four
// blocks), unless its calls are not affine; then each contraction makes the
// four updates per integral.
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "gint_math.h"
#include "c2s_table.h"

namespace gint {
namespace tq {

__host__ __device__ constexpr int nco(int l) { return (l + 1) * (l + 2) / 2; }

// component d (x y z) of cartesian function i of shell L (descending x, then y)
__host__ __device__ constexpr int comp(int L, int i, int d)
{
   int n = 0;
   for (int lx = L; lx >= 0; lx--)
      for (int ly = L - lx; ly >= 0; ly--) {
         if (n == i) return d == 0 ? lx : (d == 1 ? ly : L - lx - ly);
         n++;
      }
   return 0;
}
__host__ __device__ constexpr int cidx(int L, int lx, int lz) { return (L - lx) * (L - lx + 1) / 2 + lz; }
__host__ __device__ constexpr int first_dir(int L, int i) { return comp(L, i, 0) > 0 ? 0 : (comp(L, i, 1) > 0 ? 1 : 2); }
__host__ __device__ constexpr int shift_idx(int L, int i, int d, int by)
{
   const int lx = comp(L, i, 0) + (d == 0 ? by : 0);
   const int lz = comp(L, i, 2) + (d == 2 ? by : 0);
   return cidx(L + by, lx, lz);
}
__host__ __device__ constexpr double c2s(int l, int m, int k)
{
   constexpr double v[] = {GINT_C2S_VALUES};
   constexpr int off[4] = {GINT_C2S_OFF_0, GINT_C2S_OFF_1, GINT_C2S_OFF_2, GINT_C2S_OFF_3};
   return v[off[l] + m * nco(l) + k];
}

// layouts of one class, as in ClassPlan (plan.cpp), at compile time
template <int la, int lb, int lc, int ld>
struct Cls {
   static constexpr int LA = la + lb, LC = lc + ld;
   static constexpr int Nc = nco(la) * nco(lb) * nco(lc) * nco(ld);
   static constexpr int Ns = (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1);
   static constexpr int ncd = nco(lc) * nco(ld);
   // ket layer dd: blocks (e, c), e in [la, LA], c in [lc, LC - dd], e outer
   static constexpr int koff(int dd, int e, int c)
   {
      int off = 0;
      for (int e2 = la; e2 <= LA; e2++)
         for (int c2 = lc; c2 <= LC - dd; c2++) {
            if (e2 == e && c2 == c) return off;
            off += nco(e2) * nco(c2) * nco(dd);
         }
      return off;
   }
   static constexpr int ksize(int dd) { return koff(dd, LA + 1, lc); }
   // bra layer bb: blocks a in [la, LA - bb]; layer 0 is the last ket layer
   static constexpr int boff(int bb, int a)
   {
      int off = 0;
      for (int a2 = la; a2 <= LA - bb; a2++) {
         if (a2 == a) return off;
         off += nco(a2) * nco(bb) * ncd;
      }
      return off;
   }
   static constexpr int bsize(int bb) { return boff(bb, LA - bb + 1); }
   static constexpr int ncp = koff(0, LA + 1, lc);
   static constexpr int cap()
   {
      int c = ncp > Nc ? ncp : Nc;
      for (int dd = 1; dd <= ld; dd++) c = ksize(dd) > c ? ksize(dd) : c;
      for (int bb = 1; bb <= lb; bb++) c = bsize(bb) > c ? bsize(bb) : c;
      return c;
   }
};

// one step of HRR on a contraction's LDS row (element i at row[i]): ket
// layer dd from layer dd-1
template <int la, int lb, int lc, int ld, int dd>
__device__ __forceinline__ void ket_step_lds(const double *in, double *out, const double CD[3])
{
   using C = Cls<la, lb, lc, ld>;
#pragma unroll
   for (int e = la; e <= C::LA; e++)
#pragma unroll
      for (int c = lc; c <= C::LC - dd; c++)
#pragma unroll
         for (int ie = 0; ie < nco(e); ie++)
#pragma unroll
            for (int ic = 0; ic < nco(c); ic++)
#pragma unroll
               for (int id = 0; id < nco(dd); id++) {
                  const int dir = first_dir(dd, id);
                  const int idm = shift_idx(dd, id, dir, -1);
                  const int icp = shift_idx(c, ic, dir, 1);
                  const int i1 = C::koff(dd - 1, e, c + 1) + (ie * nco(c + 1) + icp) * nco(dd - 1) + idm;
                  const int i0 = C::koff(dd - 1, e, c) + (ie * nco(c) + ic) * nco(dd - 1) + idm;
                  const int io = C::koff(dd, e, c) + (ie * nco(c) + ic) * nco(dd) + id;
                  out[io] = in[i1] + CD[dir] * in[i0];
               }
}

// bra layer bb from layer bb-1
template <int la, int lb, int lc, int ld, int bb>
__device__ __forceinline__ void bra_step_lds(const double *in, double *out, const double AB[3])
{
   using C = Cls<la, lb, lc, ld>;
#pragma unroll
   for (int a = la; a <= C::LA - bb; a++)
#pragma unroll
      for (int ia = 0; ia < nco(a); ia++)
#pragma unroll
         for (int ib = 0; ib < nco(bb); ib++)
#pragma unroll
            for (int icd = 0; icd < C::ncd; icd++) {
               const int dir = first_dir(bb, ib);
               const int ibm = shift_idx(bb, ib, dir, -1);
               const int iap = shift_idx(a, ia, dir, 1);
               const int i1 = C::boff(bb - 1, a + 1) + (iap * nco(bb - 1) + ibm) * C::ncd + icd;
               const int i0 = C::boff(bb - 1, a) + (ia * nco(bb - 1) + ibm) * C::ncd + icd;
               const int io = C::boff(bb, a) + (ia * nco(bb) + ib) * C::ncd + icd;
               out[io] = in[i1] + AB[dir] * in[i0];
            }
}

// HRR of one contraction of one row: layer 0 in B0, (la lb|lc ld) added to
// acc; B0 and B1 are both overwritten
template <int la, int lb, int lc, int ld>
__device__ __forceinline__ void hrr_lds(double *B0, double *B1, double *acc, const double AB[3], const double CD[3])
{
   using C = Cls<la, lb, lc, ld>;
   double *bufs[2] = {B0, B1};
   int cur = 0;
   if constexpr (ld >= 1) { ket_step_lds<la, lb, lc, ld, 1>(bufs[cur], bufs[cur ^ 1], CD); cur ^= 1; }
   if constexpr (ld >= 2) { ket_step_lds<la, lb, lc, ld, 2>(bufs[cur], bufs[cur ^ 1], CD); cur ^= 1; }
   if constexpr (ld >= 3) { ket_step_lds<la, lb, lc, ld, 3>(bufs[cur], bufs[cur ^ 1], CD); cur ^= 1; }
   if constexpr (lb >= 1) { bra_step_lds<la, lb, lc, ld, 1>(bufs[cur], bufs[cur ^ 1], AB); cur ^= 1; }
   if constexpr (lb >= 2) { bra_step_lds<la, lb, lc, ld, 2>(bufs[cur], bufs[cur ^ 1], AB); cur ^= 1; }
   if constexpr (lb >= 3) { bra_step_lds<la, lb, lc, ld, 3>(bufs[cur], bufs[cur ^ 1], AB); cur ^= 1; }
   const double *res = bufs[cur];
#pragma unroll
   for (int i = 0; i < C::Nc; i++) acc[i] += res[i];
}

// one stage of cart -> sph on a contraction's LDS row: in [k][n] -> out [n][m]
template <int l, int n>
__device__ __forceinline__ void sph_stage_lds(const double *in, double *out)
{
#pragma unroll
   for (int j = 0; j < n; j++)
#pragma unroll
      for (int i = 0; i < 2 * l + 1; i++) {
         double v = 0.0;
#pragma unroll
         for (int k = 0; k < nco(l); k++)
            if (c2s(l, i, k) != 0.0) v += c2s(l, i, k) * in[k * n + j];
         out[j * (2 * l + 1) + i] = v;
      }
}

constexpr int TQ_PB = 8;     // primitive quartets per batch
constexpr int TQ_LPL = 24;   // contractions per lane in the contraction phase
constexpr int TQ_NQC = 128;  // contractions per pass, at most

struct TqwArgs {
   int nclass;
   int nq_all;          // contractions of the class per set quartet
   int nqc;             // contractions per pass (<= nq_all)
   int nla, nlb, nlc, nld;
   int sa, sb, sc, sd;  // shells of the class on the four centres
   const int *srcmap;   // class column -> column of the type's blocks
   int o_kp, o_b0, o_b1, o_acc;   // LDS offsets, doubles (coefficients at 0)
   int o_code;          // LDS byte offset of the contraction codes ia|ib<<8|ic<<16|id<<24
};

// The K updates of one (set quartet, class), summed over the set quartet's
// contractions before the atomics (the same sums as one update per integral,
// reference src/compute_KS.cpp:51-115): K_bd += -f sum_ac I P_ac, K_bc +=
// -f sum_ad I P_ad, K_ad += -f sum_bc I P_bc, K_ac += -f sum_bd I P_bd.
// I of contraction q, integral t is at I[q * WB + t]; one atomic per K
// element of the four blocks.
template <int la, int lb, int lc, int ld, int WB, typename Y>
__device__ void ks_reduced_impl(const KsGroup &ks, const Y &y, const double *I, const KTargets &kt, int lane)
{
   constexpr int NSA = 2 * la + 1, NSB = 2 * lb + 1, NSC = 2 * lc + 1, NSD = 2 * ld + 1;
   const double symm[5] = {2.0, 1.0, 0.5, 0.25, 0.0};
   const double fac = symm[(ks.flags >> 4) & 7] * kt.hf_fac;
   if (fac == 0.0) return;
   const int nla = y.nla, nlb = y.nlb, nlc = y.nlc, nld = y.nld;
   const unsigned ldac = ks.ld01 & 0xffff, ldad = ks.ld01 >> 16, ldbc = ks.ld23 & 0xffff, ldbd = ks.ld23 >> 16;
   const bool Tac = ks.flags & 1, Tad = ks.flags & 2, Tbc = ks.flags & 4, Tbd = ks.flags & 8;
   auto blk = [](bool T, unsigned ldx, int r, int c) { return T ? (size_t)r * ldx + c : (size_t)c * ldx + r; };
   auto q_of = [&](int ia, int ib, int ic, int id) { return ((ia * nlb + ib) * nlc + ic) * nld + id; };
   auto t_of = [](int isa, int isb, int isc, int isd) { return ((isa * NSB + isb) * NSC + isc) * NSD + isd; };
   for (int sp = 0; sp < kt.nspin; sp++) {
      const double *__restrict__ P = kt.P[sp];
      double *__restrict__ K = kt.K[sp];
      for (int x = lane; x < nlb * NSB * nld * NSD; x += 64) {   // K_bd
         const int ib = x / (NSB * nld * NSD), isb = (x / (nld * NSD)) % NSB, id = (x / NSD) % nld, isd = x % NSD;
         double sum = 0.0;
         for (int ia = 0; ia < nla; ia++)
            for (int ic = 0; ic < nlc; ic++) {
               const size_t o = (size_t)((long long)ks.base[0] + (long long)ia * ks.step[0] + (long long)ic * ks.step[1]);
               const double *Iq = I + q_of(ia, ib, ic, id) * WB;
#pragma unroll
               for (int isa = 0; isa < NSA; isa++)
#pragma unroll
                  for (int isc = 0; isc < NSC; isc++) sum += Iq[t_of(isa, isb, isc, isd)] * P[o + blk(Tac, ldac, isa, isc)];
            }
         const size_t ob = (size_t)((long long)ks.base[3] + (long long)ib * ks.step[6] + (long long)id * ks.step[7]);
         unsafeAtomicAdd(&K[ob + blk(Tbd, ldbd, isb, isd)], -fac * sum);
      }
      for (int x = lane; x < nlb * NSB * nlc * NSC; x += 64) {   // K_bc
         const int ib = x / (NSB * nlc * NSC), isb = (x / (nlc * NSC)) % NSB, ic = (x / NSC) % nlc, isc = x % NSC;
         double sum = 0.0;
         for (int ia = 0; ia < nla; ia++)
            for (int id = 0; id < nld; id++) {
               const size_t o = (size_t)((long long)ks.base[1] + (long long)ia * ks.step[2] + (long long)id * ks.step[3]);
               const double *Iq = I + q_of(ia, ib, ic, id) * WB;
#pragma unroll
               for (int isa = 0; isa < NSA; isa++)
#pragma unroll
                  for (int isd = 0; isd < NSD; isd++) sum += Iq[t_of(isa, isb, isc, isd)] * P[o + blk(Tad, ldad, isa, isd)];
            }
         const size_t ob = (size_t)((long long)ks.base[2] + (long long)ib * ks.step[4] + (long long)ic * ks.step[5]);
         unsafeAtomicAdd(&K[ob + blk(Tbc, ldbc, isb, isc)], -fac * sum);
      }
      for (int x = lane; x < nla * NSA * nld * NSD; x += 64) {   // K_ad
         const int ia = x / (NSA * nld * NSD), isa = (x / (nld * NSD)) % NSA, id = (x / NSD) % nld, isd = x % NSD;
         double sum = 0.0;
         for (int ib = 0; ib < nlb; ib++)
            for (int ic = 0; ic < nlc; ic++) {
               const size_t o = (size_t)((long long)ks.base[2] + (long long)ib * ks.step[4] + (long long)ic * ks.step[5]);
               const double *Iq = I + q_of(ia, ib, ic, id) * WB;
#pragma unroll
               for (int isb = 0; isb < NSB; isb++)
#pragma unroll
                  for (int isc = 0; isc < NSC; isc++) sum += Iq[t_of(isa, isb, isc, isd)] * P[o + blk(Tbc, ldbc, isb, isc)];
            }
         const size_t ob = (size_t)((long long)ks.base[1] + (long long)ia * ks.step[2] + (long long)id * ks.step[3]);
         unsafeAtomicAdd(&K[ob + blk(Tad, ldad, isa, isd)], -fac * sum);
      }
      for (int x = lane; x < nla * NSA * nlc * NSC; x += 64) {   // K_ac
         const int ia = x / (NSA * nlc * NSC), isa = (x / (nlc * NSC)) % NSA, ic = (x / NSC) % nlc, isc = x % NSC;
         double sum = 0.0;
         for (int ib = 0; ib < nlb; ib++)
            for (int id = 0; id < nld; id++) {
               const size_t o = (size_t)((long long)ks.base[3] + (long long)ib * ks.step[6] + (long long)id * ks.step[7]);
               const double *Iq = I + q_of(ia, ib, ic, id) * WB;
#pragma unroll
               for (int isb = 0; isb < NSB; isb++)
#pragma unroll
                  for (int isd = 0; isd < NSD; isd++) sum += Iq[t_of(isa, isb, isc, isd)] * P[o + blk(Tbd, ldbd, isb, isd)];
            }
         const size_t ob = (size_t)((long long)ks.base[0] + (long long)ia * ks.step[0] + (long long)ic * ks.step[1]);
         unsafeAtomicAdd(&K[ob + blk(Tac, ldac, isa, isc)], -fac * sum);
      }
   }
}

template <int la, int lb, int lc, int ld>
struct TqShape {
   using C = Cls<la, lb, lc, ld>;
   static_assert(C::ncp <= 128, "at most two columns per lane");
   static constexpr int NCOL = C::ncp > 64 ? 2 : 1;            // columns per lane
   static constexpr int LL = NCOL == 1 ? 64 / C::ncp : 1;     // lanes per column
   static constexpr int CAP = C::cap();
   // LDS row strides of one contraction (odd: the lanes' rows fall in
   // different banks), compile-time so that element offsets are immediates
   static constexpr int WB = CAP | 1, WA = C::Nc | 1;
   // contractions per lane in the contraction phase
   static constexpr int LPL = NCOL == 2 ? 12 : ((TQ_NQC + LL - 1) / LL < TQ_LPL ? (TQ_NQC + LL - 1) / LL : TQ_LPL);
};

template <int la, int lb, int lc, int ld, typename Y>
__device__ __forceinline__ void ks_reduced(const KsGroup &ks, const Y &y, const double *I, const KTargets &kt, int lane)
{
   ks_reduced_impl<la, lb, lc, ld, TqShape<la, lb, lc, ld>::WB>(ks, y, I, kt, lane);
}

template <int la, int lb, int lc, int ld>
__global__ __launch_bounds__(64) void k_tqw(int cls_k, DevTables t, BatchDev b, KTargets kt, TqwArgs y)
{
   using C = Cls<la, lb, lc, ld>;
   using S = TqShape<la, lb, lc, ld>;
   constexpr int NCP = C::ncp, LL = S::LL, NC = C::Nc, NS = C::Ns;
   extern __shared__ double lds[];
   const int lane = threadIdx.x;
   double *co = lds;               // [4 centres][4 contractions][16 primitives]
   double *kpb = lds + y.o_kp;     // [TQ_PB][nqc]
   double *B0 = lds + y.o_b0;      // [nqc][WB]
   double *B1 = lds + y.o_b1;      // [nqc][WB]
   double *ACC = lds + y.o_acc;    // [nqc][WA]
   uint32_t *code = (uint32_t *)((char *)lds + y.o_code);   // [nq_all]
   // the lane's columns (lane, and lane + 64 when the class has more than 64)
   // and its group of contractions
   constexpr int NCOL = S::NCOL;
   const int grp = NCOL == 1 ? lane / NCP : 0;
   int colv[NCOL], srcv[NCOL];
   bool cv[NCOL];
#pragma unroll
   for (int c = 0; c < NCOL; c++) {
      colv[c] = NCOL == 1 ? lane % NCP : lane + 64 * c;
      cv[c] = NCOL == 1 ? grp < LL : colv[c] < NCP;
      srcv[c] = cv[c] ? y.srcmap[colv[c]] : 0;
   }
   const bool cact = cv[0];        // lane takes part in the contraction
   const int ncpt = b.tp->ncp;
   const uint32_t nfix = min(b.fix_count[0], b.fix_cap);
   const int gate = b.tp->cls_gate[cls_k];
   const int s = y.nqc;
   {
      const int nbcd = y.nlb * y.nlc * y.nld, ncd = y.nlc * y.nld;
      for (int l = lane; l < y.nq_all; l += 64)
         code[l] = (uint32_t)(l / nbcd) | (uint32_t)((l / ncd) % y.nlb) << 8 | (uint32_t)((l / y.nld) % y.nlc) << 16 |
                   (uint32_t)(l % y.nld) << 24;
   }
   const double symm[5] = {2.0, 1.0, 0.5, 0.25, 0.0};

   for (uint32_t gi = b.g0 + blockIdx.x; gi < b.g1; gi += gridDim.x) {
      const GroupRec &g = b.groups[gi];
      const KsGroup ks = group_ks(b, *b.tp, g, cls_k);
      const int nq = (int)(ks.flags >> 16);
      if (nq == 0) continue;
      __syncthreads();
      // the set quartet's coefficients of the class's four shells
      {
         const int shv[4] = {y.sa, y.sb, y.sc, y.sd}, nlv[4] = {y.nla, y.nlb, y.nlc, y.nld};
         for (int i = lane; i < 256; i += 64) {
            const int X = i >> 6, inl = (i >> 4) & 3, ip = i & 15;
            const int np = ((g.np4 >> (4 * X)) & 15) + 1;
            co[i] = (inl < nlv[X] && ip < np) ? b.data[g.Kof[X][shv[X]] + inl * np + ip] : 0.0;
         }
      }
      for (int l0 = 0; l0 < nq; l0 += y.nqc) {
         const int nqc = min(y.nqc, nq - l0);
         for (int i = lane; i < S::WA * nqc; i += 64) ACC[i] = 0.0;
         __syncthreads();
         for (uint32_t r = g.row0; r < g.row1; r++) {
            const RowRec row = b.rows[r];
            double E[NCOL][S::LPL];
#pragma unroll
            for (int c = 0; c < NCOL; c++)
#pragma unroll
               for (int k = 0; k < S::LPL; k++) E[c][k] = 0.0;
            const double *A0 = b.ac0 + (size_t)(row.prm_off - b.p0) * ncpt;
            const uint32_t *IP = b.prm_ip + row.prm_off;
            for (uint32_t p0 = 0; p0 < row.n_prm; p0 += TQ_PB) {
               const int nb = min(TQ_PB, (int)(row.n_prm - p0));
               double av[NCOL][TQ_PB];
#pragma unroll
               for (int c = 0; c < NCOL; c++)
#pragma unroll
                  for (int q = 0; q < TQ_PB; q++) av[c][q] = (cv[c] && q < nb) ? A0[(size_t)(p0 + q) * ncpt + srcv[c]] : 0.0;
               // the batch's primitive indices, one per lane, loaded with the blocks
               const int lq8 = lane & (TQ_PB - 1);
               const uint32_t ipl = lq8 < nb ? IP[p0 + lq8] : 0u;
               __syncthreads();
               // coefficient products of the batch, once per (primitive, contraction)
               for (int q = 0; q < nb; q++) {
                  const uint32_t ip = (uint32_t)__shfl((int)ipl, q, 64);
                  for (int lq = lane; lq < nqc; lq += 64) {
                     const uint32_t cd = code[l0 + lq];
                     kpb[q * s + lq] = co[(cd & 255) * 16 + (ip & 15)] * co[64 + ((cd >> 8) & 255) * 16 + ((ip >> 4) & 15)] *
                                       co[128 + ((cd >> 16) & 255) * 16 + ((ip >> 8) & 15)] *
                                       co[192 + (cd >> 24) * 16 + ((ip >> 12) & 15)];
                  }
               }
               __syncthreads();
               if (cact) {
#pragma unroll
                  for (int k = 0; k < S::LPL; k++) {
                     const int lq = grp + k * LL;
                     if (lq < nqc) {
#pragma unroll
                        for (int q = 0; q < TQ_PB; q++)
                           if (q < nb) {
                              const double kq = kpb[q * s + lq];
#pragma unroll
                              for (int c = 0; c < NCOL; c++) E[c][k] += kq * av[c][q];
                           }
                     }
                  }
               }
            }
            // slots of this row whose screening depends on the class (rare)
            for (uint32_t e2 = 0; e2 < nfix; e2++) {
               const uint32_t p = b.fix[2 * e2];
               if (b.prm_row[p] != r) continue;
               const uint32_t ip = b.prm_ip[p];
               const double *fsrc = b.fix_ws + (size_t)e2 * (ncpt + 16);
               const bool keep = ((b.fix[2 * e2 + 1] >> 8) >> gate) & 1u;
               // a screened slot keeps its Fm[0] in (00|00) only (reference src/compute_VRR.cpp:384-385)
               double v[NCOL];
#pragma unroll
               for (int c = 0; c < NCOL; c++)
                  v[c] = !cv[c] ? 0.0
                                : (keep ? fsrc[srcv[c]] : ((la == 0 && lc == 0 && colv[c] == 0) ? fsrc[ncpt + gate] : 0.0));
               if (cact) {
#pragma unroll
                  for (int k = 0; k < S::LPL; k++) {
                     const int lq = grp + k * LL;
                     if (lq < nqc) {
                        const uint32_t cd = code[l0 + lq];
                        const double kq = co[(cd & 255) * 16 + (ip & 15)] * co[64 + ((cd >> 8) & 255) * 16 + ((ip >> 4) & 15)] *
                                          co[128 + ((cd >> 16) & 255) * 16 + ((ip >> 8) & 15)] *
                                          co[192 + (cd >> 24) * 16 + ((ip >> 12) & 15)];
#pragma unroll
                        for (int c = 0; c < NCOL; c++) E[c][k] += kq * v[c];
                     }
                  }
               }
            }
            // the row's contracted blocks, one LDS row per contraction
            if (cact) {
#pragma unroll
               for (int k = 0; k < S::LPL; k++) {
                  const int lq = grp + k * LL;
                  if (lq < nqc) {
#pragma unroll
                     for (int c = 0; c < NCOL; c++)
                        if (cv[c]) B0[lq * S::WB + colv[c]] = E[c][k];
                  }
               }
            }
            __syncthreads();
            double AB[3], CD[3];
            {
               const double *Ao = b.data + g.R[0], *Bo = b.data + g.R[1], *Co = b.data + g.R[2], *Do = b.data + g.R[3];
               const int n1 = row.n12 & 0xffff, n2 = row.n12 >> 16;
               double ABs[3], CDs[3];
               gint_pbc(Ao, Bo, t.cell, ABs);
               gint_pbc(Co, Do, t.cell, CDs);
#pragma unroll
               for (int k = 0; k < 3; k++) {
                  const double Bk = Ao[k] + ABs[k] + t.neighs[n1 * 3 + k];
                  const double Dk = Co[k] + CDs[k] + t.neighs[n2 * 3 + k];
                  AB[k] = Ao[k] - Bk;
                  CD[k] = Co[k] - Dk;
               }
            }
            for (int lq = lane; lq < nqc; lq += 64)
               hrr_lds<la, lb, lc, ld>(B0 + lq * S::WB, B1 + lq * S::WB, ACC + lq * S::WA, AB, CD);
            __syncthreads();
         }
         // cartesian -> spherical and the K updates of each contraction
         constexpr int n0 = nco(lb) * nco(lc) * nco(ld);
         constexpr int n1 = nco(lc) * nco(ld) * (2 * la + 1);
         constexpr int n2 = nco(ld) * (2 * la + 1) * (2 * lb + 1);
         constexpr int n3 = (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1);
         // all of the set quartet's contractions in one pass, with affine
         // offsets: the K updates are summed over its contractions first
         const bool out_mode = (ks.flags & KS_OUTPUT) != 0;
         const bool reduce = !(ks.flags & (KS_EXPLICIT | KS_OUTPUT)) && nq == y.nq_all && nqc == nq;
         for (int lq = lane; lq < nqc; lq += 64) {
            double *b0 = B0 + lq * S::WB, *b1 = B1 + lq * S::WB;
            sph_stage_lds<la, n0>(ACC + lq * S::WA, b0);
            sph_stage_lds<lb, n1>(b0, b1);
            sph_stage_lds<lc, n2>(b1, b0);
            sph_stage_lds<ld, n3>(b0, b1);
            if (reduce) continue;
            const int l = l0 + lq;
            const uint32_t cd = code[l];
            const long long ia = cd & 255, ib = (cd >> 8) & 255, ic = (cd >> 16) & 255, id = cd >> 24;
            if (out_mode) {
               // the integrals themselves, in the reference's TRA order
               double *O = kt.out + ((size_t)ks.base[0] | (size_t)ks.base[1] << 32);
#pragma unroll 4
               for (int tt = 0; tt < NS; tt++) {
                  const int isa = tt / ((2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1));
                  const int isb = (tt / ((2 * lc + 1) * (2 * ld + 1))) % (2 * lb + 1);
                  const int isc = (tt / (2 * ld + 1)) % (2 * lc + 1), isd = tt % (2 * ld + 1);
                  O[tra_index((int)ia, isa, (int)ib, isb, (int)ic, isc, (int)id, isd, y.nlb, y.nlc, y.nld, 2 * la + 1,
                              2 * lb + 1, 2 * lc + 1, 2 * ld + 1)] = b1[tt];
               }
               continue;
            }
            size_t off[4];
            uint32_t ld01 = ks.ld01, ld23 = ks.ld23, flags = ks.flags;
            if (ks.flags & KS_EXPLICIT) {
               const KsRec x = b.ksx[ks.expl + l];
               for (int u = 0; u < 4; u++) off[u] = x.off[u];
               ld01 = x.ld01;
               ld23 = x.ld23;
               flags = x.flags;
            } else {
               off[0] = (size_t)((long long)ks.base[0] + ia * ks.step[0] + ic * ks.step[1]);
               off[1] = (size_t)((long long)ks.base[1] + ia * ks.step[2] + id * ks.step[3]);
               off[2] = (size_t)((long long)ks.base[2] + ib * ks.step[4] + ic * ks.step[5]);
               off[3] = (size_t)((long long)ks.base[3] + ib * ks.step[6] + id * ks.step[7]);
            }
            const double fac = symm[(flags >> 4) & 7] * kt.hf_fac;
            if (fac == 0.0) continue;
            const unsigned ldac = ld01 & 0xffff, ldad = ld01 >> 16, ldbc = ld23 & 0xffff, ldbd = ld23 >> 16;
            const bool Tac = flags & 1, Tad = flags & 2, Tbc = flags & 4, Tbd = flags & 8;
            for (int sp = 0; sp < kt.nspin; sp++) {
               const double *__restrict__ P = kt.P[sp];
               double *__restrict__ K = kt.K[sp];
#pragma unroll 4
               for (int tt = 0; tt < NS; tt++) {
                  const int isa = tt / ((2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1));
                  const int isb = (tt / ((2 * lc + 1) * (2 * ld + 1))) % (2 * lb + 1);
                  const int isc = (tt / (2 * ld + 1)) % (2 * lc + 1), isd = tt % (2 * ld + 1);
                  const size_t Iac = off[0] + (Tac ? isa * ldac + isc : isc * ldac + isa);
                  const size_t Iad = off[1] + (Tad ? isa * ldad + isd : isd * ldad + isa);
                  const size_t Ibc = off[2] + (Tbc ? isb * ldbc + isc : isc * ldbc + isb);
                  const size_t Ibd = off[3] + (Tbd ? isb * ldbd + isd : isd * ldbd + isb);
                  const double iabcd = -fac * b1[tt];
                  const double pac = P[Iac], pad = P[Iad], pbc = P[Ibc], pbd = P[Ibd];
                  unsafeAtomicAdd(&K[Ibd], iabcd * pac);
                  unsafeAtomicAdd(&K[Ibc], iabcd * pad);
                  unsafeAtomicAdd(&K[Iad], iabcd * pbc);
                  unsafeAtomicAdd(&K[Iac], iabcd * pbd);
               }
            }
         }
         if (reduce) {
            __syncthreads();
            ks_reduced<la, lb, lc, ld>(ks, y, B1, kt, lane);
         }
         __syncthreads();
      }
   }
}

template <int la, int lb, int lc, int ld>
int launch_wave(int cls_k, const TypePlan &tp, const DevTables &t, const BatchDev &b, const KTargets &kt,
                const int *srcmap, hipStream_t st)
{
   using C = Cls<la, lb, lc, ld>;
   using S = TqShape<la, lb, lc, ld>;
   TqwArgs y;
   const int sh = tp.cls_sh[cls_k];
   y.sa = sh & 15;
   y.sb = (sh >> 4) & 15;
   y.sc = (sh >> 8) & 15;
   y.sd = (sh >> 12) & 15;
   y.nla = tp.nl[0][y.sa];
   y.nlb = tp.nl[1][y.sb];
   y.nlc = tp.nl[2][y.sc];
   y.nld = tp.nl[3][y.sd];
   y.nclass = tp.nclass;
   y.nq_all = tp.cls_nl[cls_k];
   y.nqc = std::min(std::min(y.nq_all, S::LPL * S::LL), TQ_NQC);
   y.srcmap = srcmap;
   const int nqc = y.nqc;
   y.o_kp = 256;
   y.o_b0 = y.o_kp + TQ_PB * nqc;
   y.o_b1 = y.o_b0 + S::WB * nqc;
   y.o_acc = y.o_b1 + S::WB * nqc;
   y.o_code = 8 * (y.o_acc + S::WA * nqc);
   const size_t lds_bytes = (size_t)y.o_code + 4 * (size_t)y.nq_all;
   if (y.nq_all < 1 || y.nq_all > 256 || lds_bytes > 64 * 1024) return -6;
   const unsigned ng = b.g1 - b.g0;
   if (ng == 0) return 0;
   const unsigned grid = std::min(ng, 1u << 16);
   hipLaunchKernelGGL((k_tqw<la, lb, lc, ld>), dim3(grid), dim3(64), lds_bytes, st, cls_k, t, b, kt, y);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

// instantiated in kernels_tq_part.hip, GINT_TQ_NPART translation units
// compiled in parallel; called by the dispatcher in kernels_tq.hip
template <int la, int lb, int lc, int ld>
int launch_one(int cls_k, const TypePlan &tp, const DevTables &t, const BatchDev &b, const KTargets &kt,
               const int *srcmap, hipStream_t st);

}  // namespace tq

// the classes with a wavefront-per-set-quartet kernel, in GINT_TQ_NPART
// groups of similar compile cost (one translation unit each)
#define GINT_TQ_NPART 8
#define GINT_TQ_CLASSES_0(X) X(1, 0, 2, 2) X(2, 0, 2, 0) X(0, 2, 2, 0) X(1, 2, 0, 1) X(0, 0, 1, 2) X(1, 1, 0, 1) X(0, 2, 0, 0) X(0, 0, 0, 1)
#define GINT_TQ_CLASSES_1(X) X(2, 2, 1, 0) X(2, 2, 0, 0) X(1, 1, 2, 0) X(0, 1, 1, 2) X(2, 0, 0, 1) X(0, 1, 1, 1) X(0, 0, 0, 2) X(0, 0, 0, 0)
#define GINT_TQ_CLASSES_2(X) X(2, 1, 2, 0) X(0, 0, 2, 2) X(2, 1, 1, 0) X(0, 2, 0, 2) X(0, 1, 2, 0) X(0, 1, 0, 2) X(1, 0, 0, 1)
#define GINT_TQ_CLASSES_3(X) X(2, 0, 2, 1) X(1, 0, 2, 1) X(2, 1, 0, 1) X(2, 1, 0, 0) X(1, 1, 1, 0) X(2, 0, 0, 0) X(0, 1, 1, 0)
#define GINT_TQ_CLASSES_4(X) X(1, 2, 2, 0) X(1, 0, 1, 2) X(1, 2, 1, 0) X(0, 0, 2, 1) X(1, 0, 1, 1) X(0, 0, 2, 0) X(0, 1, 0, 1)
#define GINT_TQ_CLASSES_5(X) X(2, 0, 1, 2) X(0, 1, 2, 1) X(1, 1, 1, 1) X(2, 0, 1, 0) X(1, 0, 0, 2) X(1, 1, 0, 0) X(1, 0, 0, 0)
#define GINT_TQ_CLASSES_6(X) X(2, 2, 0, 1) X(2, 0, 0, 2) X(0, 2, 1, 1) X(1, 0, 2, 0) X(0, 2, 1, 0) X(1, 0, 1, 0) X(0, 0, 1, 0)
#define GINT_TQ_CLASSES_7(X) X(0, 1, 2, 2) X(2, 0, 1, 1) X(1, 1, 0, 2) X(1, 2, 0, 0) X(0, 2, 0, 1) X(0, 0, 1, 1) X(0, 1, 0, 0)
#define GINT_TQ_CLASSES(X) GINT_TQ_CLASSES_0(X) GINT_TQ_CLASSES_1(X) GINT_TQ_CLASSES_2(X) GINT_TQ_CLASSES_3(X) GINT_TQ_CLASSES_4(X) GINT_TQ_CLASSES_5(X) GINT_TQ_CLASSES_6(X) GINT_TQ_CLASSES_7(X)

}  // namespace gint
