// Per-class recurrence plan: workspace layouts of the vertical (Obara-Saika)
// and horizontal (Head-Gordon-Pople) recurrences for (la lb|lc ld).
//
// Same recurrence choices as reference src/plan.cpp:56-146 (rec_add): the
// bra VRR builds a from the (0|0)^(m) ladder, the ket VRR builds c with the
// cross term into a-1, HRR2 moves b onto a, HRR1 moves d onto c, the
// cartesian direction is the first non-zero one (x, y, z).  The layout is
// this library's own: instead of a list of 19-int ops, the kernels walk
// fixed level loops over blocks whose offsets are computed here.
#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

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

   // VRR blocks: ket level c needs a >= max(0, la - (LC - c)); level 0 needs
   // the whole bra ladder 0..LA.  Block (a,c) carries m = 0..L-a-c.
   p.amin[0] = 0;
   for (int c = 1; c <= p.LC; c++) p.amin[c] = std::max(0, la - (p.LC - c));
   off = 0;
   for (int c = 0; c <= p.LC; c++)
      for (int a = p.amin[c]; a <= p.LA; a++) {
         p.voff[a][c] = off;
         off += (p.L - a - c + 1) * nco(a) * nco(c);
      }
   p.vws = off;

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

}  // namespace gint
