// Plans of the recurrences.
//
// ClassPlan: horizontal (Head-Gordon-Pople) recurrence layers and
// cartesian -> spherical sizes of one class (la lb|lc ld).  HRR2 moves b
// onto a, HRR1 moves d onto c, the cartesian direction is the first non-zero
// one (x, y, z), as in reference src/plan.cpp:56-146 (rec_add).  Instead of
// a list of 19-int ops the kernels walk fixed level loops over blocks whose
// offsets are computed here.
//
// TypePlan: the union of the (e0|f0) blocks that the classes of one quartet
// type read, the Hermite sizes of the McMurchie-Davidson path that produces
// them, and the per-class bookkeeping (gates, contraction layout, checksum
// multipliers).
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "device.hpp"

namespace gint {

// VRR FLOPs of the reference plan for one class: the recursion of reference
// src/plan.cpp:56-146 from the (e0|f0) targets at m = 0, with the m-ranges
// of each (a,c) block collapsed to [min m, max m] (src/plan.cpp:320-338) and
// weights per element VRR1 3, VRR2 8, VRR3 3, VRR4 8, VRR5 11, VRR6 6.
static int ref_vrr_flops(int la, int lb, int lc, int ld)
{
   std::map<std::pair<int, int>, std::pair<int, int>> mr;  // (a,c) -> (min m, max m)
   std::set<std::tuple<int, int, int>> seen;
   std::vector<std::tuple<int, int, int>> stack;
   for (int e = la; e <= la + lb; e++)
      for (int f = lc; f <= lc + ld; f++) stack.emplace_back(e, f, 0);
   while (!stack.empty()) {
      auto [a, c, m] = stack.back();
      stack.pop_back();
      if (a + c == 0 || !seen.insert({a, c, m}).second) continue;
      auto it = mr.find({a, c});
      if (it == mr.end()) mr[{a, c}] = {m, m};
      else it->second = {std::min(it->second.first, m), std::max(it->second.second, m)};
      if (c == 0) {
         stack.emplace_back(a - 1, 0, m);
         stack.emplace_back(a - 1, 0, m + 1);
         if (a > 1) {
            stack.emplace_back(a - 2, 0, m);
            stack.emplace_back(a - 2, 0, m + 1);
         }
      } else {
         stack.emplace_back(a, c - 1, m);
         stack.emplace_back(a, c - 1, m + 1);
         if (c > 1) {
            stack.emplace_back(a, c - 2, m);
            stack.emplace_back(a, c - 2, m + 1);
         }
         if (a > 0) stack.emplace_back(a - 1, c - 1, m + 1);
      }
   }
   long fl = 0;
   for (auto &kv : mr) {
      const int a = kv.first.first, c = kv.first.second;
      int w;
      if (c == 0) w = (a == 1) ? 3 : 8;
      else if (a == 0) w = (c == 1) ? 3 : 8;
      else w = (c == 1) ? 6 : 11;
      fl += (long)w * (kv.second.second - kv.second.first + 1) * nco(a) * nco(c);
   }
   return (int)fl;
}

ClassPlan make_plan(int la, int lb, int lc, int ld)
{
   ClassPlan p;
   std::memset(&p, 0, sizeof p);
   p.la = la; p.lb = lb; p.lc = lc; p.ld = ld;
   p.LA = la + lb;
   p.LC = lc + ld;
   p.L = p.LA + p.LC;
   p.Nc = nco(la) * nco(lb) * nco(lc) * nco(ld);
   p.Ns = nsph(la) * nsph(lb) * nsph(lc) * nsph(ld);

   // (e0|f0) blocks that feed the HRR, e outer, f inner
   int off = 0;
   for (int e = la; e <= p.LA; e++)
      for (int f = lc; f <= p.LC; f++) {
         p.cpoff[e][f] = off;
         off += nco(e) * nco(f);
      }
   p.ncp = off;

   // HRR layers
   int cap = std::max(p.ncp, p.Nc);
   for (int dd = 0; dd <= ld; dd++) {
      off = 0;
      for (int e = la; e <= p.LA; e++)
         for (int c = lc; c <= p.LC - dd; c++) {
            p.koff[dd][e][c] = off;
            off += nco(e) * nco(c) * nco(dd);
         }
      p.ksize[dd] = off;
      cap = std::max(cap, off);
   }
   const int ncd = nco(lc) * nco(ld);
   for (int bb = 0; bb <= lb; bb++) {
      off = 0;
      for (int a = la; a <= p.LA - bb; a++) {
         p.boff[bb][a] = off;
         off += nco(a) * nco(bb) * ncd;
      }
      p.bsize[bb] = off;
      cap = std::max(cap, off);
   }
   p.ycap = cap;
   p.hrr_ws = 3 * cap + p.Nc;

   p.fm_flops = 52 + 211 * (p.L + 1);
   p.vrr_flops = ref_vrr_flops(la, lb, lc, ld);
   int hf = 0;
   for (int dd = 1; dd <= ld; dd++) hf += 2 * p.ksize[dd];
   for (int bb = 1; bb <= lb; bb++) hf += 2 * p.bsize[bb];
   p.hrr_flops = hf;
   {
      int dims[4] = {nco(la), nco(lb), nco(lc), nco(ld)};
      const int ls[4] = {la, lb, lc, ld};
      int sf = 0;
      for (int s = 0; s < 4; s++) {
         const int k = dims[0], n = dims[1] * dims[2] * dims[3], m = nsph(ls[s]);
         sf += 2 * m * n * k;
         dims[0] = dims[1]; dims[1] = dims[2]; dims[2] = dims[3]; dims[3] = m;
      }
      p.sph_flops = sf;
   }
   return p;
}

// odd multiplier of a class in the survivor checksum (the oracle uses the
// same function)
uint64_t class_mult(int code)
{
   uint64_t z = 0xC1A55E5ULL + (uint64_t)code * 0x9E3779B97F4A7C15ULL;
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   z = z ^ (z >> 31);
   return z | 1ull;
}

// FLOPs of the Hermite recursion R^(n)_tuv, n = L-1..0, from the G_m ladder:
// an element whose recursion direction has component >= 2 costs 3 (one
// product with (t-1), one fused product-add), otherwise 1; plus the (-2 rho)^n
// scaling (2 per order) and the cell accumulation of the n = 0 layer.
static long long herm_flops(int L)
{
   long long f = 2LL * (L + 1);
   for (int n = L - 1; n >= 0; n--)
      for (int N = 1; N <= L - n; N++)
         for (int t = N; t >= 0; t--)
            for (int u = N - t; u >= 0; u--) {
               const int v = N - t - u;
               const int c = t > 0 ? t : (u > 0 ? u : v);
               f += c >= 2 ? 3 : 1;
            }
   return f + nherm(L);
}

void make_type_plan(TypePlan &tp)
{
   TypePlan p;
   std::memset(&p, 0, sizeof p);
   for (int t = 0; t < 4; t++) {
      p.nsh[t] = tp.nsh[t];
      for (int s = 0; s < tp.nsh[t]; s++) {
         p.l[t][s] = tp.l[t][s];
         p.nl[t][s] = tp.nl[t][s];
      }
   }
   int lmin[4], lmax[4];
   for (int t = 0; t < 4; t++) {
      lmin[t] = 99;
      lmax[t] = -1;
      for (int s = 0; s < p.nsh[t]; s++) {
         lmin[t] = std::min(lmin[t], p.l[t][s]);
         lmax[t] = std::max(lmax[t], p.l[t][s]);
      }
   }
   p.LA0 = lmin[0];
   p.LA = lmax[0] + lmax[1];
   p.LC0 = lmin[2];
   p.LC = lmax[2] + lmax[3];
   p.L = p.LA + p.LC;
   int off = 0;
   for (int e = p.LA0; e <= p.LA; e++)
      for (int f = p.LC0; f <= p.LC; f++) {
         p.cpoff[e][f] = off;
         off += nco(e) * nco(f);
      }
   p.ncp = off;
   off = 0;
   for (int f = p.LC0; f <= p.LC; f++) {
      p.kcoff[f] = off;
      off += nco(f);
   }
   p.nkc = off;
   p.nbh = nherm(p.LA);
   p.nh = nherm(p.L);
   p.has00 = (p.LA0 == 0 && p.LC0 == 0) ? 1 : 0;

   std::set<int> Lset;
   int k = 0, eoff = 0;
   long long ref = 0;
   for (int sa = 0; sa < p.nsh[0]; sa++)
      for (int sb = 0; sb < p.nsh[1]; sb++)
         for (int sc = 0; sc < p.nsh[2]; sc++)
            for (int sd = 0; sd < p.nsh[3]; sd++, k++) {
               const int la = p.l[0][sa], lb = p.l[1][sb], lc = p.l[2][sc], ld = p.l[3][sd];
               const ClassPlan cp = make_plan(la, lb, lc, ld);
               p.cls_code[k] = class_code(la, lb, lc, ld);
               p.cls_sh[k] = sa | sb << 4 | sc << 8 | sd << 12;
               p.cls_nl[k] = p.nl[0][sa] * p.nl[1][sb] * p.nl[2][sc] * p.nl[3][sd];
               p.cls_ncp[k] = cp.ncp;
               p.cls_eoff[k] = eoff;
               eoff += p.cls_nl[k] * cp.ncp;
               p.cls_mult[k] = class_mult(p.cls_code[k]);
               p.mult_all += p.cls_mult[k];
               Lset.insert(cp.L);
               ref += cp.fm_flops + cp.vrr_flops;
            }
   p.nclass = k;
   p.ebytes = eoff;
   int co = 0;
   for (int t = 0; t < 4; t++)
      for (int s2 = 0; s2 < p.nsh[t]; s2++) {
         p.co_off[t][s2] = co;
         co += p.nl[t][s2] * MAXNP;
      }
   p.nco_tot = co;
   for (int t = 0; t < 4; t++) {
      int f = 0;
      for (int s2 = 0; s2 < p.nsh[t]; s2++) {
         p.fun0[t][s2] = f;
         f += p.nl[t][s2] * nsph(p.l[t][s2]);
      }
      p.nfun[t] = f;
   }
   p.nL = 0;
   for (int L : Lset) p.Ls[p.nL++] = L;
   for (k = 0; k < p.nclass; k++) {
      const int c = p.cls_code[k];
      const int L = c / 64 + (c / 16) % 4 + (c / 4) % 4 + c % 4;
      for (int i = 0; i < p.nL; i++)
         if (p.Ls[i] == L) p.cls_gate[k] = i;
   }
   p.ref_slot_flops = ref;
   p.surv_flops = 52 + 211LL * (p.L + 1) + herm_flops(p.L);
   // Hermite -> cartesian: E coefficients (3 directions, 3 FLOPs per entry
   // on each side), then 3 FLOPs per term of the two contraction stages
   long long pf = 0;
   for (int e = 0; e <= p.LA; e++) pf += 3LL * 3 * (e + 1);
   for (int f = 0; f <= p.LC; f++) pf += 3LL * 3 * (f + 1);
   for (int N = 0; N <= p.LA; N++)
      for (int t = N; t >= 0; t--)
         for (int u = N - t; u >= 0; u--)
            for (int f = p.LC0; f <= p.LC; f++)
               for (int fx = f; fx >= 0; fx--)
                  for (int fy = f - fx; fy >= 0; fy--) pf += 3LL * (fx + 1) * (fy + 1) * (f - fx - fy + 1);
   for (int e = p.LA0; e <= p.LA; e++)
      for (int ex = e; ex >= 0; ex--)
         for (int ey = e - ex; ey >= 0; ey--) pf += 3LL * p.nkc * (ex + 1) * (ey + 1) * (e - ex - ey + 1);
   p.prim_flops = pf;
   tp = p;
}

void make_type_items(const TypePlan &tp, std::vector<uint32_t> &e1, std::vector<uint32_t> &e2)
{
   // cartesian components of shell L in order: lx descending, then ly descending
   auto comps = [](int L) {
      std::vector<std::array<int, 3>> v;
      for (int lx = L; lx >= 0; lx--)
         for (int ly = L - lx; ly >= 0; ly--) v.push_back({lx, ly, L - lx - ly});
      return v;
   };
   std::vector<std::array<int, 3>> ket, bra;
   for (int f = tp.LC0; f <= tp.LC; f++)
      for (auto &c : comps(f)) ket.push_back(c);
   for (int N = 0; N <= tp.LA; N++)
      for (auto &h : comps(N)) bra.push_back(h);
   // stage 1, W[h][kc] = sum over the ket E of R: ket component outer, so the
   // lanes of a wavefront share their loop bounds (fx, fy, fz)
   e1.clear();
   const int nkc = (int)ket.size();
   for (int kc = 0; kc < nkc; kc++)
      for (int h = 0; h < (int)bra.size(); h++) {
         const auto &t = bra[h], &c = ket[kc];
         e1.push_back((uint32_t)t[0] | (uint32_t)t[1] << 3 | (uint32_t)t[2] << 6 | (uint32_t)c[0] << 9 |
                      (uint32_t)c[1] << 12 | (uint32_t)c[2] << 15 | (uint32_t)(h * nkc + kc) << 18);
      }
   // stage 2, (e0|f0) = sum over the bra E of W: bra component outer for the
   // same reason; each item names its destination j in the type's blocks
   e2.clear();
   for (int e = tp.LA0; e <= tp.LA; e++) {
      const auto ce = comps(e);
      for (int ie = 0; ie < nco(e); ie++)
         for (int f = tp.LC0; f <= tp.LC; f++)
            for (int jl = 0; jl < nco(f); jl++) {
               const int j = tp.cpoff[e][f] + ie * nco(f) + jl;
               e2.push_back((uint32_t)ce[ie][0] | (uint32_t)ce[ie][1] << 3 | (uint32_t)ce[ie][2] << 6 |
                            (uint32_t)(tp.kcoff[f] + jl) << 9 | (uint32_t)j << 16);
            }
   }
}

}  // namespace gint
