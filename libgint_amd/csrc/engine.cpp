// Host engine behind the libgint_* C-ABI.  See engine.hpp.
//
// Built with g++ -fopenmp (libgomp) so that the OpenMP-collective calls of
// the protocol (init, set_P*, set_K*, get_K*) synchronise with a caller
// built by gfortran/gcc, like CP2K.
#include "engine.hpp"

#include <hip/hip_runtime_api.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gint_math.h"

namespace gint {

void fatal(const char *what, long long code)
{
   std::fprintf(stderr, "libgint: %s (%lld)\n", what, code);
   std::fflush(stderr);
   std::abort();
}

#define HIPCHECK(x)                                                   \
   do {                                                               \
      hipError_t e_ = (x);                                            \
      if (e_ != hipSuccess) fatal(#x ": " __FILE__, (int)e_);         \
   } while (0)

// ---------------------------------------------------------------------------
// basis data store
// ---------------------------------------------------------------------------
static uint64_t hash_doubles(const double *v, int n)
{
   uint64_t h = 0x9E3779B97F4A7C15ULL ^ (uint64_t)n;
   for (int i = 0; i < n; i++) {
      uint64_t x;
      std::memcpy(&x, &v[i], 8);
      h ^= x + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
      h = (h ^ (h >> 31)) * 0xbf58476d1ce4e5b9ULL;
   }
   return h;
}

uint32_t DataStore::add(const double *v, int n)
{
   const uint64_t h = hash_doubles(v, n);
   auto range = index_.equal_range(h);
   for (auto it = range.first; it != range.second; ++it) {
      const uint32_t off = it->second.first;
      if (it->second.second == n && std::memcmp(&buf_[off], v, sizeof(double) * (size_t)n) == 0) return off;
   }
   const uint32_t off = (uint32_t)buf_.size();
   buf_.insert(buf_.end(), v, v + n);
   index_.emplace(h, std::make_pair(off, n));
   version_++;
   return off;
}

template <class T>
static void free_vec(std::vector<T> &v)
{
   std::vector<T>().swap(v);
}

void TypeQueue::release()
{
   free_vec(rows);
   free_vec(prm_ip);
   free_vec(prm_row);
   free_vec(groups);
   free_vec(ksg);
   free_vec(kss);
   free_vec(ksx);
}

void TypeDev::release()
{
   void *p[9] = {rows, groups, prm_ip, prm_row, ksg, ksx, kss, gmag, gflag};
   for (void *x : p)
      if (x) (void)hipFree(x);
   gmag = nullptr;
   gflag = nullptr;
   rows = nullptr;
   groups = nullptr;
   prm_ip = prm_row = nullptr;
   ksg = nullptr;
   kss = nullptr;
   ksx = nullptr;
   for (auto &c : cap) c = 0;
   n_rows = n_groups = n_prm = n_ksx = n_ksg = n_kss = 0;
   max_group_bytes = 0;
   chunks.clear();
}

// ---------------------------------------------------------------------------
// team-shared objects: streams per OpenMP thread (reference
// src/libGint.cpp:26,37-60) and the device P / K of the current step
// ---------------------------------------------------------------------------
static std::vector<hipStream_t> g_streams;
static double *g_P[2] = {nullptr, nullptr}, *g_K[2] = {nullptr, nullptr};
static size_t g_FP = 0, g_FP_cap = 0;
static int g_nspin = 0;

static void shared_alloc(size_t n, int nspin)
{
   if (n > g_FP_cap || nspin > g_nspin) {
      for (int s = 0; s < 2; s++) {
         if (g_P[s]) HIPCHECK(hipFree(g_P[s]));
         if (g_K[s]) HIPCHECK(hipFree(g_K[s]));
         g_P[s] = g_K[s] = nullptr;
      }
      for (int s = 0; s < nspin; s++) {
         HIPCHECK(hipMalloc((void **)&g_P[s], sizeof(double) * std::max<size_t>(n, 1)));
         HIPCHECK(hipMalloc((void **)&g_K[s], sizeof(double) * std::max<size_t>(n, 1)));
      }
      g_FP_cap = n;
   }
   g_FP = n;
   g_nspin = nspin;
}

Engine::Engine()
{
   my_thr_ = omp_get_thread_num();
   nomp_ = omp_get_num_threads();
   for (int i = 0; i < 18; i++) cell_[i] = (i % 9) % 4 == 0 ? 1.0 : 0.0;
   neighs_.assign(3, 0.0);
   set_max_mem(4096);
}

Engine::~Engine()
{
   if (stream_) (void)hipStreamSynchronize((hipStream_t)stream_);
   drop_graph();
   if (own_stream_ && stream_) (void)hipStreamDestroy((hipStream_t)stream_);
   for (auto &d : stream_dev_) d.release();
   for (auto &d : plan_dev_) d.release();
   for (auto &ti : types_) {
      if (ti.d_plan) (void)hipFree(ti.d_plan);
      if (ti.d_e1) (void)hipFree(ti.d_e1);
      if (ti.d_e2) (void)hipFree(ti.d_e2);
      if (ti.d_cps) (void)hipFree(ti.d_cps);
      if (ti.d_lays) (void)hipFree(ti.d_lays);
      if (ti.d_srcmaps) (void)hipFree(ti.d_srcmaps);
      if (ti.d_eco_info) (void)hipFree(ti.d_eco_info);
      if (ti.d_eco_items) (void)hipFree(ti.d_eco_items);
   }
   void *bufs[] = {d_data_, d_C0_, d_grid_, d_bw_, d_ftab_, d_cell_, d_neighs_, d_work_,
                   d_fix_, d_fix_ws_, d_scratch_, d_stats_, d_fixcnt_, d_out_, d_mix_};
   for (void *b : bufs)
      if (b) (void)hipFree(b);
   for (void *e : ev_pool_) (void)hipEventDestroy((hipEvent_t)e);
   for (void *e : ev_join_)
      if (e) (void)hipEventDestroy((hipEvent_t)e);
   for (void *q : side_)
      if (q) (void)hipStreamDestroy((hipStream_t)q);
}

void Engine::init()
{
   if ((int)g_streams.size() < omp_get_num_threads()) {
#pragma omp barrier
#pragma omp master
      g_streams.resize(omp_get_num_threads(), nullptr);
#pragma omp barrier
   }
   if (my_thr_ >= (int)g_streams.size()) fatal("handle created outside the team that calls init", my_thr_);
   if (!g_streams[my_thr_]) HIPCHECK(hipStreamCreateWithFlags(&g_streams[my_thr_], hipStreamNonBlocking));
   stream_ = g_streams[my_thr_];
   ensure_side();
   // the device address of the no-allocation scratch (once per device)
   noalloc_scratch_ = noalloc_scratch();
}

void Engine::ensure_side()
{
   for (int i = 0; i < NSIDE; i++)
      if (!side_[i]) HIPCHECK(hipStreamCreateWithFlags((hipStream_t *)&side_[i], hipStreamNonBlocking));
   for (int i = 0; i <= NSIDE; i++)
      if (!ev_join_[i]) HIPCHECK(hipEventCreateWithFlags((hipEvent_t *)&ev_join_[i], hipEventDisableTiming));
}

// Any table of whole orders is accepted (reference src/libGint.cpp:62-100
// copies whatever it is given); a quartet type whose total angular momentum
// exceeds the table's highest order is refused when it is used.
void Engine::set_potential_truncated(double R_cut, const double *C0, int ld_C0, int C0_size)
{
   if (ld_C0 < TCG_NCOEF || ld_C0 % TCG_NCOEF != 0 || (long long)C0_size < (long long)ld_C0 * TCG_NPATCH)
      fatal("set_Potential_Truncated: C0 must hold 207 patches of ld_C0 = (orders) x 105 coefficients", ld_C0);
   truncated_ = true;
   R_cut_ = R_cut;
   ld_C0_ = ld_C0;
   nder_ = ld_C0 / TCG_NCOEF - 1;
   C0_.assign(C0, C0 + (size_t)ld_C0 * TCG_NPATCH);
   tables_dirty_ = true;
}

// reference src/libGint.cpp:108-115: the budget is shared by the team
void Engine::set_max_mem(int max_mem_MiB)
{
   budget_bytes_ = (size_t)std::max(max_mem_MiB, 64) * 1024 * 1024 / (size_t)std::max(nomp_, 1);
}

void Engine::set_cell(const double *h, const double *hinv)
{
   std::memcpy(cell_, h, 9 * sizeof(double));
   std::memcpy(cell_ + 9, hinv, 9 * sizeof(double));
   tables_dirty_ = true;
}

void Engine::set_neighs(const double *g, int n)
{
   if (n < 1 || n > MAXNG) fatal("set_neighs: number of lattice vectors must be 1..256", n);
   dispatch_all();
   neighs_.assign(g, g + 3 * n);
   tables_dirty_ = true;
}

// reference src/libGint.cpp:117-148
void Engine::set_atom(int i, const double *R, const double *Z, int np)
{
   if (i < 0 || i > 0xffff) fatal("set_Atom: set index out of 0..65535", i);
   if (np < 1 || np > MAXNP) fatal("set_Atom: 1..16 primitives per set", np);
   if (i >= (int)sets_.size()) sets_.resize(i + 1);
   SetInfo &s = sets_[i];
   s.R = store_.add(R, 3);
   s.Z = store_.add(Z, np);
   s.np = np;
   s.l.clear();
   s.nl.clear();
   s.K.clear();
   s.sig = -1;
   last_type_ = -1;
}

void Engine::set_atom_L(int i, int l, int nl, const double *K)
{
   if (i < 0 || i >= (int)sets_.size() || sets_[i].np == 0) fatal("set_Atom_L before set_Atom", i);
   if (l < 0 || l > MAXL) fatal("set_Atom_L: l must be 0..3", l);
   if (nl < 1 || nl > MAXNL) fatal("set_Atom_L: 1..4 contractions per shell", nl);
   SetInfo &s = sets_[i];
   for (int x : s.l)
      if (x == l) fatal("set_Atom_L: one shell per l and set", l);
   s.l.push_back(l);
   s.nl.push_back(nl);
   s.K.push_back(store_.add(K, nl * s.np));
   s.sig = -1;
   last_type_ = -1;
}

int Engine::set_sig(int i)
{
   SetInfo &s = sets_[i];
   if (s.sig >= 0) return s.sig;
   if (s.l.empty()) fatal("set without shells (set_Atom_L) used in add_shell", i);
   auto key = std::make_pair(s.l, s.nl);
   auto it = sig_ids_.find(key);
   if (it == sig_ids_.end()) it = sig_ids_.emplace(key, (int)sig_ids_.size()).first;
   s.sig = it->second;
   return s.sig;
}

// The matrix-core contraction of a type: its classes with CqLayout::eco,
// one work item per 64 columns x 64 contractions of each, and the row
// length of the contraction buffer.
static void make_eco_items(TypeInfo &ti)
{
   const TypePlan &tp = ti.plan;
   ti.eco_items.clear();
   ti.eco_row = 0;
   ti.eco_info.assign(2 * MAXTC, 0);
   for (int k = 0; k < tp.nclass; k++) {
      const CqLayout &y = ti.lays[k];
      ti.eco_info[k] = y.o_src;
      ti.eco_info[MAXTC + k] = y.eco_off;
      const int cols = tp.cls_ncp[k], nq = tp.cls_nl[k];
      if (y.eco_off >= 0) ti.eco_row = std::max(ti.eco_row, (size_t)y.eco_off + (size_t)nq * cols);
      if (!y.eco) continue;
      for (int jt0 = 0; jt0 * 16 < cols; jt0 += 4)
         for (int qb = 0; qb * 64 < nq; qb++) ti.eco_items.push_back(EcoItem{(uint16_t)k, (uint16_t)jt0, (uint16_t)qb, 0});
   }
   // items in the order of the first type column they read: the classes
   // share the row's blocks (a block (e,f) serves every class with
   // la <= e <= la+lb, lc <= f <= lc+ld), so items that run together read
   // the same columns of a row that is larger than the caches
   auto first_col = [&](const EcoItem &it) { return ti.srcmaps[ti.lays[it.k].o_src + it.jt0 * 16]; };
   std::stable_sort(ti.eco_items.begin(), ti.eco_items.end(), [&](const EcoItem &a, const EcoItem &b) {
      const int ca = first_col(a), cb = first_col(b);
      return ca < cb || (ca == cb && (a.k < b.k || (a.k == b.k && a.qb < b.qb)));
   });
}

int Engine::type_of(int i, int j, int k, int l)
{
   const std::array<int, 4> sets = {i, j, k, l};
   if (sets == last_sets_ && last_type_ >= 0) return last_type_;
   const std::array<int, 4> sig = {set_sig(i), set_sig(j), set_sig(k), set_sig(l)};
   auto it = type_ids_.find(sig);
   if (it == type_ids_.end()) {
      TypeInfo ti;
      ti.sig = sig;
      std::memset(&ti.plan, 0, sizeof ti.plan);
      const SetInfo *S[4] = {&sets_[i], &sets_[j], &sets_[k], &sets_[l]};
      for (int t = 0; t < 4; t++) {
         ti.plan.nsh[t] = (int)S[t]->l.size();
         if (ti.plan.nsh[t] > MAXSH) fatal("more than 4 shells in a set", ti.plan.nsh[t]);
         for (int s = 0; s < ti.plan.nsh[t]; s++) {
            ti.plan.l[t][s] = S[t]->l[s];
            ti.plan.nl[t][s] = S[t]->nl[s];
         }
      }
      make_type_plan(ti.plan);
      if (truncated_ && ti.plan.L > nder_)
         fatal("the truncated-Coulomb table has fewer derivative orders than this quartet's total angular momentum",
               ti.plan.L);
      make_type_items(ti.plan, ti.e1, ti.e2);
      if (!fm_r_lds_bytes(ti.plan)) fatal("quartet type exceeds the on-chip budget", ti.plan.L);
      for (int c = 0; c < ti.plan.nclass; c++) {
         const int code = ti.plan.cls_code[c];
         ti.cplans.push_back(make_plan(code / 64, (code / 16) % 4, (code / 4) % 4, code % 4));
      }
      make_cq_plan(ti.plan, ti.cplans, ti.lays, ti.srcmaps, &ti.cq_lds, &ti.cq_gws);
      for (const CqLayout &y : ti.lays) ti.cq_wide = ti.cq_wide || (y.kron && !y.skip);
      make_eco_items(ti);
      types_.push_back(ti);
      q_.emplace_back();
      it = type_ids_.emplace(sig, (int)types_.size() - 1).first;
   }
   last_sets_ = sets;
   last_type_ = it->second;
   return last_type_;
}

// ---------------------------------------------------------------------------
// P and K (collective, reference src/libGint.cpp:436-653)
// ---------------------------------------------------------------------------
void Engine::set_P(const double *P, int n)
{
#pragma omp single
   {
      shared_alloc((size_t)n, 1);
      HIPCHECK(hipMemcpy(g_P[0], P, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemset(g_K[0], 0, sizeof(double) * (size_t)n));
   }
   nspin_ = 1;
   FP_size_ = (size_t)n;
   d_P_[0] = g_P[0];
   d_K_[0] = g_K[0];
}

void Engine::set_P(const double *Pa, const double *Pb, int n)
{
#pragma omp single
   {
      shared_alloc((size_t)n, 2);
      HIPCHECK(hipMemcpy(g_P[0], Pa, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(g_P[1], Pb, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemset(g_K[0], 0, sizeof(double) * (size_t)n));
      HIPCHECK(hipMemset(g_K[1], 0, sizeof(double) * (size_t)n));
   }
   nspin_ = 2;
   FP_size_ = (size_t)n;
   for (int s = 0; s < 2; s++) {
      d_P_[s] = g_P[s];
      d_K_[s] = g_K[s];
   }
}

void Engine::set_K(const double *K, int n)
{
   if (nspin_ != 1 || (size_t)n != FP_size_) fatal("set_K: call set_P first, with the same size", n);
#pragma omp single
   HIPCHECK(hipMemcpy(g_K[0], K, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
}

void Engine::set_K(const double *Ka, const double *Kb, int n)
{
   if (nspin_ != 2 || (size_t)n != FP_size_) fatal("set_K_polarized: call set_P_polarized first", n);
#pragma omp single
   {
      HIPCHECK(hipMemcpy(g_K[0], Ka, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(g_K[1], Kb, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
   }
}

void Engine::get_K(double *K)
{
   if (nspin_ != 1) fatal("get_K: state is not closed-shell", nspin_);
   dispatch_all();
   synchronize();
#pragma omp barrier
#pragma omp single
   HIPCHECK(hipMemcpy(K, g_K[0], sizeof(double) * FP_size_, hipMemcpyDeviceToHost));
}

void Engine::get_K(double *Ka, double *Kb)
{
   if (nspin_ != 2) fatal("get_K_polarized: state is not spin-polarised", nspin_);
   dispatch_all();
   synchronize();
#pragma omp barrier
#pragma omp single
   {
      HIPCHECK(hipMemcpy(Ka, g_K[0], sizeof(double) * FP_size_, hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(Kb, g_K[1], sizeof(double) * FP_size_, hipMemcpyDeviceToHost));
   }
}

void Engine::synchronize()
{
   if (!stream_) return;
   HIPCHECK(hipStreamSynchronize((hipStream_t)stream_));
   harvest_events();
   if (d_fixcnt_) {
      // slots whose screening differs between the classes of a type but
      // found the fix-up list full were kept by every class: an error of at
      // most their G_0 ~ 1e-12 times the contraction coefficients in K
      uint32_t c[2];
      HIPCHECK(hipMemcpy(c, d_fixcnt_, sizeof c, hipMemcpyDeviceToHost));
      if (c[1]) {
         if (!fix_overflow_)
            std::fprintf(stderr, "libgint: %u slots with class-dependent screening beyond the fix-up list of %zu "
                         "were kept by every class\n", c[1], fix_cap_);
         fix_overflow_ += c[1];
         uint32_t z = 0;
         HIPCHECK(hipMemcpy(d_fixcnt_ + 1, &z, sizeof z, hipMemcpyHostToDevice));
      }
   }
}

// ---------------------------------------------------------------------------
// work submission, reference src/libGint.cpp:151-414
// ---------------------------------------------------------------------------
void Engine::add_prm(int ipa, int ipb, int ipc, int ipd)
{
   if ((unsigned)ipa >= MAXNP || (unsigned)ipb >= MAXNP || (unsigned)ipc >= MAXNP || (unsigned)ipd >= MAXNP)
      fatal("add_prm: primitive index out of range", ipa);
   prm_tmp_.push_back((uint32_t)ipa | ((uint32_t)ipb << 4) | ((uint32_t)ipc << 8) | ((uint32_t)ipd << 12));
}

void Engine::add_shell(int i, int j, int k, int l, int n1, int n2)
{
   if (prm_tmp_.empty()) return;   // reference src/libGint.cpp:159
   const int ns = (int)sets_.size();
   if (i < 0 || j < 0 || k < 0 || l < 0 || i >= ns || j >= ns || k >= ns || l >= ns) fatal("add_shell: unknown set", i);
   const int Ng = (int)neighs_.size() / 3;
   if (n1 < 0 || n2 < 0 || n1 >= Ng || n2 >= Ng) fatal("add_shell: cell index beyond set_neighs", n1 > n2 ? n1 : n2);
   const SetInfo *S[4] = {&sets_[i], &sets_[j], &sets_[k], &sets_[l]};
   for (uint32_t ip : prm_tmp_)
      for (int t = 0; t < 4; t++)
         if ((int)((ip >> (4 * t)) & 15) >= S[t]->np) fatal("add_prm: primitive index beyond the set's primitives", t);
   const int type = type_of(i, j, k, l);
   TypeQueue &Q = q_[type];
   if (cur_type_ < 0) {
      cur_type_ = type;
      cur_sets_ = {i, j, k, l};
      GroupRec g;
      std::memset(&g, 0, sizeof g);
      g.row0 = (uint32_t)Q.rows.size();
      for (int t = 0; t < 4; t++) {
         g.R[t] = S[t]->R;
         g.Z[t] = S[t]->Z;
         for (size_t s = 0; s < S[t]->K.size(); s++) g.Kof[t][s] = S[t]->K[s];
      }
      g.np4 = (uint32_t)(S[0]->np - 1) | (uint32_t)(S[1]->np - 1) << 4 | (uint32_t)(S[2]->np - 1) << 8 |
              (uint32_t)(S[3]->np - 1) << 12;
      g.set_ij = (uint32_t)i | (uint32_t)j << 16;
      g.set_kl = (uint32_t)k | (uint32_t)l << 16;
      g.ks = 0;
      g.ksmode = 0;
      Q.groups.push_back(g);
      pend_bytes_ += sizeof(GroupRec) + sizeof(KsGroup) * (size_t)types_[type].plan.nclass;
   } else if (cur_sets_ != std::array<int, 4>{i, j, k, l}) {
      fatal("add_shell: rows of a different set quartet before libgint_add_set", i);
   }
   if (Q.prm_ip.size() + prm_tmp_.size() > 0xffffffffull) fatal("more than 2^32 primitive quartets of one type", 0);
   RowRec r;
   r.prm_off = (uint32_t)Q.prm_ip.size();
   r.n_prm = (uint32_t)prm_tmp_.size();
   r.group = (uint32_t)Q.groups.size() - 1;
   r.n12 = (uint32_t)n1 | (uint32_t)n2 << 16;
   const uint32_t row_idx = (uint32_t)Q.rows.size();
   Q.rows.push_back(r);
   Q.prm_ip.insert(Q.prm_ip.end(), prm_tmp_.begin(), prm_tmp_.end());
   Q.prm_row.insert(Q.prm_row.end(), prm_tmp_.size(), row_idx);
   pend_bytes_ += sizeof(RowRec) + 8 * prm_tmp_.size();
   prm_tmp_.clear();
}

void Engine::add_qrt(int la, int lb, int lc, int ld, int nla, int nlb, int nlc, int nld)
{
   if (la < 0 || la > MAXL || lb < 0 || lb > MAXL || lc < 0 || lc > MAXL || ld < 0 || ld > MAXL)
      fatal("add_qrt: l must be 0..3", la);
   if (nla < 1 || nlb < 1 || nlc < 1 || nld < 1 || nla > MAXNL || nlb > MAXNL || nlc > MAXNL || nld > MAXNL)
      fatal("add_qrt: 1..4 contractions per shell", nla);
}

// symmetry factor codes, reference src/libGint.cpp:262-267
static uint32_t symm_code(double f)
{
   if (f < 0.1) return 4;
   if (f < 0.3) return 3;
   if (f < 0.75) return 2;
   if (f < 1.5) return 1;
   return 0;
}

void Engine::add_qrtt(double symm_fac, int la, int lb, int lc, int ld, int inla, int inlb, int inlc, int inld,
                      int ld_ac, int ld_ad, int ld_bc, int ld_bd, unsigned off_ac, unsigned off_ad, unsigned off_bc,
                      unsigned off_bd, bool Tac, bool Tad, bool Tbc, bool Tbd)
{
   // the position of a call among its class's calls, not inla..inld, selects
   // its contraction (reference src/libGint.cpp:250-280)
   (void)inla; (void)inlb; (void)inlc; (void)inld;
   if (la < 0 || la > MAXL || lb < 0 || lb > MAXL || lc < 0 || lc > MAXL || ld < 0 || ld > MAXL)
      fatal("add_qrtt: l must be 0..3", la);
   if (ld_ac < 0 || ld_ad < 0 || ld_bc < 0 || ld_bd < 0 || ld_ac > 0xffff || ld_ad > 0xffff || ld_bc > 0xffff ||
       ld_bd > 0xffff)
      fatal("add_qrtt: leading dimension out of 0..65535", ld_ac);
   const int c = class_code(la, lb, lc, ld);
   if (pend_[c].empty()) touched_.push_back(c);
   KsRec e;
   e.off[0] = off_ac;
   e.off[1] = off_ad;
   e.off[2] = off_bc;
   e.off[3] = off_bd;
   e.ld01 = (uint32_t)ld_ac | (uint32_t)ld_ad << 16;
   e.ld23 = (uint32_t)ld_bc | (uint32_t)ld_bd << 16;
   e.flags = (uint32_t)Tac | (uint32_t)Tad << 1 | (uint32_t)Tbc << 2 | (uint32_t)Tbd << 3 | symm_code(symm_fac) << 4;
   e.pad = 0;
   pend_[c].push_back(e);
   n_q_++;
}

// The add_qrtt calls of one (set quartet, class) as a KsGroup: offsets
// affine in (inla, inlb, inlc, inld) (the contraction order of the calls,
// row-major) with shared leading dimensions and flags, else an explicit list.
KsGroup Engine::compress(std::vector<KsRec> &calls, const TypePlan &tp, int k, std::vector<KsRec> &ksx)
{
   KsGroup g;
   std::memset(&g, 0, sizeof g);
   const size_t n = calls.size();
   if (n == 0) return g;
   const int sh = tp.cls_sh[k];
   const int nla = tp.nl[0][sh & 15], nlb = tp.nl[1][(sh >> 4) & 15], nlc = tp.nl[2][(sh >> 8) & 15],
             nld = tp.nl[3][(sh >> 12) & 15];
   const size_t nq = (size_t)nla * nlb * nlc * nld;
   if (n > nq) fatal("add_qrtt: more calls than contraction combinations of the class", (long long)n);
   // highest element of P / K the calls reach (checked against the matrix
   // size before any kernel reads or updates them)
   {
      const int code = tp.cls_code[k];
      const uint64_t ns[4] = {(uint64_t)nsph(code / 64), (uint64_t)nsph((code / 16) % 4), (uint64_t)nsph((code / 4) % 4),
                              (uint64_t)nsph(code % 4)};
      const int rc[4][2] = {{0, 2}, {0, 3}, {1, 2}, {1, 3}};   // ac ad bc bd: row / column centre
      for (const KsRec &c : calls)
         for (int q = 0; q < 4; q++) {
            const uint64_t ld = q < 2 ? (q == 0 ? c.ld01 & 0xffff : c.ld01 >> 16) : (q == 2 ? c.ld23 & 0xffff : c.ld23 >> 16);
            const uint64_t nr = ns[rc[q][0]], nc = ns[rc[q][1]];
            const uint64_t top = c.off[q] + ((c.flags >> q) & 1 ? (nr - 1) * ld + (nc - 1) : (nc - 1) * ld + (nr - 1));
            if (top > max_index_) max_index_ = top;
         }
   }
   const KsRec &c0 = calls[0];
   for (int q = 0; q < 4; q++) g.base[q] = c0.off[q];
   g.ld01 = c0.ld01;
   g.ld23 = c0.ld23;
   long long st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
   auto diff = [&](size_t l, int q) { return (long long)calls[l].off[q] - (long long)c0.off[q]; };
   if (nld > 1 && n > 1) { st[3] = diff(1, 1); st[7] = diff(1, 3); }
   if (nlc > 1 && n > (size_t)nld) { st[1] = diff(nld, 0); st[5] = diff(nld, 2); }
   if (nlb > 1 && n > (size_t)(nlc * nld)) { st[4] = diff(nlc * nld, 2); st[6] = diff(nlc * nld, 3); }
   if (nla > 1 && n > (size_t)(nlb * nlc * nld)) { st[0] = diff(nlb * nlc * nld, 0); st[2] = diff(nlb * nlc * nld, 1); }
   bool ok = true;
   for (long long s : st)
      if (s < INT32_MIN || s > INT32_MAX) ok = false;
   for (size_t l = 0; ok && l < n; l++) {
      const long long ia = (long long)(l / (nlb * nlc * nld)), ib = (long long)((l / (nlc * nld)) % nlb);
      const long long ic = (long long)((l / nld) % nlc), id = (long long)(l % nld);
      const KsRec &c = calls[l];
      ok = c.ld01 == c0.ld01 && c.ld23 == c0.ld23 && c.flags == c0.flags &&
           (long long)c.off[0] == (long long)c0.off[0] + ia * st[0] + ic * st[1] &&
           (long long)c.off[1] == (long long)c0.off[1] + ia * st[2] + id * st[3] &&
           (long long)c.off[2] == (long long)c0.off[2] + ib * st[4] + ic * st[5] &&
           (long long)c.off[3] == (long long)c0.off[3] + ib * st[6] + id * st[7];
   }
   if (ok) {
      for (int q = 0; q < 8; q++) g.step[q] = (int32_t)st[q];
      g.flags = c0.flags | (uint32_t)n << 16;
   } else {
      g.flags = KS_EXPLICIT | (uint32_t)n << 16;
      if (ksx.size() + n > 0xffffffffull) fatal("too many explicit add_qrtt records", 0);
      g.expl = (uint32_t)ksx.size();
      ksx.insert(ksx.end(), calls.begin(), calls.end());
      pend_bytes_ += sizeof(KsRec) * n;
   }
   return g;
}

// Whether the per-class groups g[0..nclass) of a set quartet are the
// set-level layout of one KsSet (every class complete, one leading dimension,
// transpose and symmetry code per pair block, offsets following the shell
// order of the sets), and that record.
bool Engine::set_level(const KsGroup *g, const TypePlan &tp, KsSet &s) const
{
   const uint32_t f0 = g[0].flags & 0x7fu;
   for (int k = 0; k < tp.nclass; k++) {
      if ((g[k].flags & KS_EXPLICIT) || ks_count(g[k].flags) != (uint32_t)tp.cls_nl[k]) return false;
      if ((g[k].flags & 0x7fu) != f0 || g[k].ld01 != g[0].ld01 || g[k].ld23 != g[0].ld23) return false;
   }
   // class 0 takes shell 0 of every set: its first call is function (0,0) of the pair blocks
   for (int q = 0; q < 4; q++) s.base[q] = g[0].base[q];
   s.ld01 = g[0].ld01;
   s.ld23 = g[0].ld23;
   s.flags = f0;
   s.pad = 0;
   for (int k = 0; k < tp.nclass; k++) {
      const KsGroup d = ks_from_set(s, tp, k);
      for (int q = 0; q < 4; q++)
         if (d.base[q] != g[k].base[q]) return false;
      const int sh = tp.cls_sh[k];
      const bool mul[4] = {tp.nl[0][sh & 15] > 1, tp.nl[1][(sh >> 4) & 15] > 1, tp.nl[2][(sh >> 8) & 15] > 1,
                           tp.nl[3][(sh >> 12) & 15] > 1};
      // step q moves contraction index of centre ctr[q] (a c a d b c b d)
      const int ctr[8] = {0, 2, 0, 3, 1, 2, 1, 3};
      for (int q = 0; q < 8; q++)
         if (mul[ctr[q]] && d.step[q] != g[k].step[q]) return false;
   }
   return true;
}

void Engine::add_set_blocks(double symm_fac, const int ld[4], const unsigned off[4], const bool T[4])
{
   if (!touched_.empty()) fatal("add_set_blocks after add_qrtt calls of the same set quartet", 0);
   if (cur_type_ < 0) return;   // no row with primitives: nothing to add, as for add_qrtt
   if (have_blocks_) fatal("add_set_blocks twice for one set quartet", 0);
   for (int q = 0; q < 4; q++)
      if (ld[q] < 0 || ld[q] > 0xffff) fatal("add_set_blocks: leading dimension out of 0..65535", ld[q]);
   KsSet &s = blocks_;
   for (int q = 0; q < 4; q++) s.base[q] = off[q];
   s.ld01 = (uint32_t)ld[0] | (uint32_t)ld[1] << 16;
   s.ld23 = (uint32_t)ld[2] | (uint32_t)ld[3] << 16;
   s.flags = (uint32_t)T[0] | (uint32_t)T[1] << 1 | (uint32_t)T[2] << 2 | (uint32_t)T[3] << 3 | symm_code(symm_fac) << 4;
   s.pad = 0;
   const TypePlan &tp = types_[cur_type_].plan;
   // the last element of each pair block (rows of a / b, columns of c / d)
   const int rc[4][2] = {{0, 2}, {0, 3}, {1, 2}, {1, 3}};
   for (int q = 0; q < 4; q++) {
      const uint64_t r = (uint64_t)tp.nfun[rc[q][0]] - 1, c = (uint64_t)tp.nfun[rc[q][1]] - 1;
      const uint64_t top = (uint64_t)off[q] + (T[q] ? r * (uint64_t)ld[q] + c : c * (uint64_t)ld[q] + r);
      if (top > max_index_) max_index_ = top;
   }
   for (int k = 0; k < tp.nclass; k++) n_q_ += (uint64_t)tp.cls_nl[k];
   have_blocks_ = true;
}

void Engine::set_output(double *out, uint64_t n)
{
   dispatch_all();
   synchronize();
   if (d_out_) HIPCHECK(hipFree(d_out_));
   d_out_ = nullptr;
   h_out_ = out;
   out_n_ = out ? n : 0;
   if (out_n_) {
      HIPCHECK(hipMalloc((void **)&d_out_, sizeof(double) * out_n_));
      HIPCHECK(hipMemset(d_out_, 0, sizeof(double) * out_n_));
   }
}

void Engine::add_tra(int la, int lb, int lc, int ld, uint64_t dest)
{
   if (!d_out_) fatal("add_tra without an output array (libgint_amd_set_output)", 0);
   if (la < 0 || la > MAXL || lb < 0 || lb > MAXL || lc < 0 || lc > MAXL || ld < 0 || ld > MAXL)
      fatal("add_tra: l must be 0..3", la);
   if (cur_type_ < 0) return;   // no primitive survived: the integrals are zero, as the array starts
   const TypePlan &tp = types_[cur_type_].plan;
   const int c = class_code(la, lb, lc, ld);
   int k = -1;
   for (int x = 0; x < tp.nclass; x++)
      if (tp.cls_code[x] == c) k = x;
   if (k < 0) fatal("add_tra: the set quartet has no such class", c);
   const uint64_t n = (uint64_t)tp.cls_nl[k] * (uint64_t)(nsph(la) * nsph(lb) * nsph(lc) * nsph(ld));
   if (dest + n > out_n_) fatal("add_tra: the class's integrals reach beyond the output array", (long long)dest);
   for (auto &e : pend_out_)
      if (e.first == c) fatal("add_tra: class given twice in one set quartet", c);
   pend_out_.push_back({c, dest});
}

void Engine::get_output()
{
   if (!d_out_) fatal("get_output without an output array", 0);
   dispatch_all();
   synchronize();
   HIPCHECK(hipMemcpy(h_out_, d_out_, sizeof(double) * out_n_, hipMemcpyDeviceToHost));
}

void Engine::add_set()
{
   if (cur_type_ >= 0) {
      TypeQueue &Q = q_[cur_type_];
      const TypePlan &tp = types_[cur_type_].plan;
      GroupRec &gr = Q.groups.back();
      gr.row1 = (uint32_t)Q.rows.size();
      const size_t per_class = sizeof(KsGroup) * (size_t)tp.nclass;
      if (have_blocks_) {
         gr.ks = (uint32_t)Q.kss.size();
         gr.ksmode = KS_SETLEVEL;
         Q.kss.push_back(blocks_);
         pend_bytes_ = pend_bytes_ - per_class + sizeof(KsSet);
      } else {
         const size_t first = Q.ksg.size();
         const size_t nx = Q.ksx.size();
         for (int k = 0; k < tp.nclass; k++) Q.ksg.push_back(compress(pend_[tp.cls_code[k]], tp, k, Q.ksx));
         for (const auto &e : pend_out_) {
            for (int k = 0; k < tp.nclass; k++) {
               if (tp.cls_code[k] != e.first) continue;
               KsGroup &g = Q.ksg[first + k];
               if (ks_count(g.flags)) fatal("add_tra and add_qrtt calls for one class of a set quartet", e.first);
               std::memset(&g, 0, sizeof g);
               g.base[0] = (uint32_t)(e.second & 0xffffffffu);
               g.base[1] = (uint32_t)(e.second >> 32);
               g.flags = KS_OUTPUT | (uint32_t)tp.cls_nl[k] << 16;
            }
         }
         KsSet s;
         if (Q.ksx.size() == nx && pend_out_.empty() && set_level(&Q.ksg[first], tp, s)) {
            Q.ksg.resize(first);
            gr.ks = (uint32_t)Q.kss.size();
            gr.ksmode = KS_SETLEVEL;
            Q.kss.push_back(s);
            pend_bytes_ = pend_bytes_ - per_class + sizeof(KsSet);
         } else {
            if (first > 0xffffffffull - (size_t)tp.nclass) fatal("more than 2^32 class records of one quartet type", 0);
            gr.ks = (uint32_t)first;
            gr.ksmode = 0;
         }
      }
   } else if (have_blocks_) {
      fatal("add_set_blocks without a set quartet", 0);
   }
   have_blocks_ = false;
   pend_out_.clear();
   // calls of classes without integrals in this set quartet add nothing
   for (int c : touched_) pend_[c].clear();
   touched_.clear();
   cur_type_ = -1;
   if (!recording_ && pend_bytes_ > budget_bytes_ / 2) dispatch_all();
}

void Engine::dispatch_all()
{
   if (cur_type_ >= 0 || !touched_.empty()) fatal("dispatch with an open set: call add_set first", cur_type_);
   bool any = false;
   for (auto &Q : q_) any = any || !Q.empty();
   if (!any || recording_) return;
   if (!d_P_[0]) fatal("integrals dispatched before set_P", 0);
   synchronize();   // the previous batch no longer reads the device arrays
   upload(q_, stream_dev_, false);
   for (auto &Q : q_) {
      Q.rows.clear();
      Q.prm_ip.clear();
      Q.prm_row.clear();
      Q.groups.clear();
      Q.ksg.clear();
      Q.kss.clear();
      Q.ksx.clear();
   }
   pend_bytes_ = 0;
   const uint64_t mi = max_index_;
   max_index_ = 0;
   execute(stream_dev_, mi);
}

// ---------------------------------------------------------------------------
// plan mode: the recorded work stays on the device
// ---------------------------------------------------------------------------
void Engine::plan_begin()
{
   dispatch_all();
   recording_ = true;
}

void Engine::plan_end()
{
   if (!recording_) fatal("plan_end without plan_begin", 0);
   if (cur_type_ >= 0) fatal("plan_end with an open set: call add_set first", cur_type_);
   recording_ = false;
   synchronize();
   drop_graph();
   for (auto &d : plan_dev_) d.release();
   plan_dev_.clear();
   upload(q_, plan_dev_, true);
   plan_tuned_ = false;
   mixed_calibrated_ = false;
   // tables and basis data resident now, so that an execution needs no
   // allocation (plan_execute_noalloc)
   ensure_tables();
   upload_data();
   pend_bytes_ = 0;
   plan_max_index_ = max_index_;
   max_index_ = 0;
}

void Engine::plan_execute()
{
   if (recording_) fatal("plan_execute while recording", 0);
   if (!d_P_[0]) fatal("plan_execute before set_P", 0);
   dispatch_all();
   RunBufs w{};
   KTargets kt{};
   if (!prepare_exec(plan_dev_, plan_max_index_, w, kt, FP_size_)) return;
   // (not while an integral-output array is set: its classes write there)
   if (autotune_on_ && !plan_tuned_ && !d_out_) autotune(w, kt);
   if (debug_ & 24) force_kron(plan_dev_, (debug_ & 16) != 0);
   mixed_mode_ = 0;
   if (mixed_tau_ > 0.0 && !d_out_) {
      ensure_mixed();
      mixed_mode_ = mixed_calibrated_ ? 2 : 1;
   }
   // per-launch events, device counters and host waits stay eager; so does
   // a handle of a team of several threads: while one thread captured its
   // stream, the others make allocating and synchronous calls of the
   // protocol (plan_end, set_P, get_K) that a capture in progress forbids
   if (graph_on_ && nomp_ == 1 && !timing_on_ && !stats_on_ && (debug_ & 6) == 0)
      replay_graph(w, kt);
   else
      launch_all(plan_dev_, w, kt);
   if (mixed_mode_ == 1) mixed_calibrated_ = true;
   mixed_mode_ = 0;
}

// device arrays of the mixed-precision route (plans)
void Engine::ensure_mixed()
{
   if (!d_mix_) {
      HIPCHECK(hipMalloc((void **)&d_mix_, 2 * sizeof(unsigned long long)));
      HIPCHECK(hipMemset(d_mix_, 0, 2 * sizeof(unsigned long long)));
   }
   for (auto &D : plan_dev_) {
      if (!D.n_groups || D.gmag) continue;
      HIPCHECK(hipMalloc((void **)&D.gmag, sizeof(double) * D.n_groups));
      HIPCHECK(hipMalloc((void **)&D.gflag, sizeof(uint32_t) * D.n_groups));
      mixed_calibrated_ = false;
   }
}

void Engine::mixed_info(double out[4])
{
   synchronize();
   out[0] = mixed_tau_;
   out[1] = mixed_calibrated_ ? 1.0 : 0.0;
   unsigned long long c[2] = {0, 0};
   if (d_mix_) HIPCHECK(hipMemcpy(c, d_mix_, sizeof c, hipMemcpyDeviceToHost));
   out[2] = (double)c[1];
   uint64_t ng = 0;
   for (auto &D : plan_dev_) ng += D.n_groups;
   out[3] = (double)ng;
}

void Engine::plan_clear()
{
   synchronize();
   drop_graph();
   for (auto &d : plan_dev_) d.release();
   plan_dev_.clear();
}

void Engine::queue_info(uint64_t out[8])
{
   for (int i = 0; i < 8; i++) out[i] = 0;
   for (size_t ty = 0; ty < q_.size(); ty++) {
      const TypeQueue &Q = q_[ty];
      if (Q.empty()) continue;
      out[0]++;
      out[1] += Q.groups.size();
      out[2] += Q.rows.size();
      out[3] += Q.prm_ip.size();
      for (const KsGroup &g : Q.ksg) {
         if (ks_count(g.flags)) out[4]++;
         if (g.flags & KS_EXPLICIT) out[5]++;
      }
      out[4] += Q.kss.size() * (uint64_t)types_[ty].plan.nclass;
      out[6] += Q.kss.size();
      for (const GroupRec &g : Q.groups) {
         size_t gp = 0;
         for (uint32_t r = g.row0; r < g.row1; r++) gp += Q.rows[r].n_prm;
         out[7] = std::max<uint64_t>(out[7], (uint64_t)gp * sizeof(double) * (uint64_t)types_[ty].plan.ncp);
      }
   }
}

int Engine::queue_types(uint64_t *out, int max_types)
{
   int n = 0;
   for (size_t ty = 0; ty < q_.size(); ty++) {
      const TypeQueue &Q = q_[ty];
      if (Q.empty()) continue;
      if (n < max_types) {
         const TypePlan &tp = types_[ty].plan;
         uint64_t *o = out + 8 * (size_t)n;
         o[0] = (uint64_t)tp.L;
         o[1] = (uint64_t)tp.ncp;
         o[2] = (uint64_t)tp.nclass;
         o[3] = Q.groups.size();
         o[4] = Q.rows.size();
         o[5] = Q.prm_ip.size();
         uint64_t nq = 0;
         for (int k = 0; k < tp.nclass; k++) nq += (uint64_t)tp.cls_nl[k];
         o[6] = nq;   // contracted quartets per set quartet
         o[7] = (uint64_t)(tp.nsh[0] | tp.nsh[1] << 4 | tp.nsh[2] << 8 | tp.nsh[3] << 12);
      }
      n++;
   }
   return n;
}

void Engine::plan_info(uint64_t out[6])
{
   uint64_t bytes = 0, chunks = 0, groups = 0, rows = 0, prm = 0, nt = 0;
   for (auto &d : plan_dev_) {
      if (d.n_groups == 0) continue;
      nt++;
      bytes += d.n_rows * sizeof(RowRec) + d.n_groups * sizeof(GroupRec) + d.n_prm * 8 + d.n_ksg * sizeof(KsGroup) +
               d.n_kss * sizeof(KsSet) + d.n_ksx * sizeof(KsRec);
      chunks += d.chunks.size();
      groups += d.n_groups;
      rows += d.n_rows;
      prm += d.n_prm;
   }
   out[0] = bytes;
   out[1] = nt;
   out[2] = chunks;
   out[3] = groups;
   out[4] = rows;
   out[5] = prm;
}

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------
void Engine::ensure_tables()
{
   if (!tables_dirty_) return;
   if (!d_grid_) {
      std::vector<int> grid(2 * TCG_GRID * TCG_GRID);
      std::vector<double> bw(4 * TCG_NPATCH);
      gint_fill_patch_tables(grid.data(), grid.data() + TCG_GRID * TCG_GRID, bw.data());
      std::vector<double> ft(121 * 28);
      boys_table(ft.data());
      HIPCHECK(hipMalloc((void **)&d_grid_, sizeof(int) * grid.size()));
      HIPCHECK(hipMalloc((void **)&d_bw_, sizeof(double) * bw.size()));
      HIPCHECK(hipMalloc((void **)&d_ftab_, sizeof(double) * ft.size()));
      HIPCHECK(hipMalloc((void **)&d_cell_, sizeof(double) * 18));
      HIPCHECK(hipMemcpy(d_grid_, grid.data(), sizeof(int) * grid.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(d_bw_, bw.data(), sizeof(double) * bw.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(d_ftab_, ft.data(), sizeof(double) * ft.size(), hipMemcpyHostToDevice));
   }
   if (truncated_) {
      if (d_C0_) HIPCHECK(hipFree(d_C0_));
      HIPCHECK(hipMalloc((void **)&d_C0_, sizeof(double) * C0_.size()));
      HIPCHECK(hipMemcpy(d_C0_, C0_.data(), sizeof(double) * C0_.size(), hipMemcpyHostToDevice));
   }
   HIPCHECK(hipMemcpy(d_cell_, cell_, sizeof(double) * 18, hipMemcpyHostToDevice));
   if (neighs_.size() > d_neighs_cap_) {
      if (d_neighs_) HIPCHECK(hipFree(d_neighs_));
      HIPCHECK(hipMalloc((void **)&d_neighs_, sizeof(double) * neighs_.size()));
      d_neighs_cap_ = neighs_.size();
   }
   HIPCHECK(hipMemcpy(d_neighs_, neighs_.data(), sizeof(double) * neighs_.size(), hipMemcpyHostToDevice));
   tables_dirty_ = false;
}

void Engine::upload_data()
{
   if (data_version_dev_ == store_.version()) return;
   const std::vector<double> &d = store_.data();
   if (d.size() > d_data_cap_) {
      if (d_data_) HIPCHECK(hipFree(d_data_));
      const size_t cap = std::max(d.size() * 2, (size_t)4096);
      HIPCHECK(hipMalloc((void **)&d_data_, sizeof(double) * cap));
      d_data_cap_ = cap;
   }
   HIPCHECK(hipMemcpy(d_data_, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice));
   data_version_dev_ = store_.version();
}

template <class T>
static void dev_put(T *&ptr, size_t &cap, const std::vector<T> &v)
{
   if (v.size() > cap || !ptr) {
      if (ptr) HIPCHECK(hipFree(ptr));
      const size_t c = std::max<size_t>(v.size() + v.size() / 2, 16);
      HIPCHECK(hipMalloc((void **)&ptr, sizeof(T) * c));
      cap = c;
   }
   if (!v.empty()) HIPCHECK(hipMemcpy(ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
}

// Copy the queues to the device and split each type's work into chunks
// whose primitive blocks and contracted rows fit the work buffer.
void Engine::upload(std::vector<TypeQueue> &qs, std::vector<TypeDev> &devs, bool release_host)
{
   const size_t work = budget_bytes_ / 2;
   if (devs.size() < qs.size()) devs.resize(qs.size());
   for (size_t ty = 0; ty < qs.size(); ty++) {
      TypeQueue &Q = qs[ty];
      TypeDev &D = devs[ty];
      D.type = (int)ty;
      D.chunks.clear();
      D.chunks_na.clear();
      D.n_groups = D.n_rows = D.n_prm = D.n_ksx = 0;
      if (Q.empty()) continue;
      TypeInfo &ti = types_[ty];
      const TypePlan &tp = ti.plan;
      if (!ti.d_plan) {
         HIPCHECK(hipMalloc((void **)&ti.d_plan, sizeof(TypePlan)));
         HIPCHECK(hipMemcpy(ti.d_plan, &tp, sizeof(TypePlan), hipMemcpyHostToDevice));
         HIPCHECK(hipMalloc((void **)&ti.d_e1, sizeof(uint32_t) * ti.e1.size()));
         HIPCHECK(hipMemcpy(ti.d_e1, ti.e1.data(), sizeof(uint32_t) * ti.e1.size(), hipMemcpyHostToDevice));
         HIPCHECK(hipMalloc((void **)&ti.d_e2, sizeof(uint32_t) * ti.e2.size()));
         HIPCHECK(hipMemcpy(ti.d_e2, ti.e2.data(), sizeof(uint32_t) * ti.e2.size(), hipMemcpyHostToDevice));
         HIPCHECK(hipMalloc((void **)&ti.d_cps, sizeof(ClassPlan) * ti.cplans.size()));
         HIPCHECK(hipMemcpy(ti.d_cps, ti.cplans.data(), sizeof(ClassPlan) * ti.cplans.size(), hipMemcpyHostToDevice));
         HIPCHECK(hipMalloc((void **)&ti.d_lays, sizeof(CqLayout) * ti.lays.size()));
         HIPCHECK(hipMemcpy(ti.d_lays, ti.lays.data(), sizeof(CqLayout) * ti.lays.size(), hipMemcpyHostToDevice));
         HIPCHECK(hipMalloc((void **)&ti.d_srcmaps, sizeof(int) * ti.srcmaps.size()));
         HIPCHECK(hipMemcpy(ti.d_srcmaps, ti.srcmaps.data(), sizeof(int) * ti.srcmaps.size(), hipMemcpyHostToDevice));
         HIPCHECK(hipMalloc((void **)&ti.d_eco_info, sizeof(int) * ti.eco_info.size()));
         HIPCHECK(hipMemcpy(ti.d_eco_info, ti.eco_info.data(), sizeof(int) * ti.eco_info.size(), hipMemcpyHostToDevice));
         if (!ti.eco_items.empty()) {
            HIPCHECK(hipMalloc((void **)&ti.d_eco_items, sizeof(EcoItem) * ti.eco_items.size()));
            HIPCHECK(hipMemcpy(ti.d_eco_items, ti.eco_items.data(), sizeof(EcoItem) * ti.eco_items.size(),
                               hipMemcpyHostToDevice));
         }
      }
      dev_put(D.rows, D.cap[0], Q.rows);
      dev_put(D.groups, D.cap[1], Q.groups);
      dev_put(D.prm_ip, D.cap[2], Q.prm_ip);
      dev_put(D.prm_row, D.cap[3], Q.prm_row);
      dev_put(D.ksg, D.cap[4], Q.ksg);
      dev_put(D.ksx, D.cap[5], Q.ksx);
      dev_put(D.kss, D.cap[6], Q.kss);
      D.n_groups = Q.groups.size();
      D.n_rows = Q.rows.size();
      D.n_prm = Q.prm_ip.size();
      D.n_ksx = Q.ksx.size();
      D.n_ksg = Q.ksg.size();
      D.n_kss = Q.kss.size();
      D.max_group_bytes = 0;
      // the K layout of class k of group gi, from whichever record holds it
      auto gks = [&](size_t gi, int k) {
         const GroupRec &g = Q.groups[gi];
         return (g.ksmode & KS_SETLEVEL) ? ks_from_set(Q.kss[g.ks], tp, k) : Q.ksg[(size_t)g.ks + k];
      };
      D.class_used.assign(tp.nclass, 0);
      for (size_t gi = 0; gi < D.n_groups; gi++) {
         if (Q.groups[gi].ksmode & KS_SETLEVEL) {
            for (int k = 0; k < tp.nclass; k++) D.class_used[k] = 1;
            continue;
         }
         for (int k = 0; k < tp.nclass; k++)
            if (ks_count(gks(gi, k).flags)) D.class_used[k] = 1;
      }
      // chunks of whole groups whose primitive blocks fit `wbytes`; a set
      // quartet larger than the allocating path's buffer gets a chunk of its
      // own, for which that buffer grows (ensure_work)
      std::vector<size_t> gbytes(D.n_groups);
      for (size_t gi = 0; gi < D.n_groups; gi++) {
         const GroupRec &g = Q.groups[gi];
         size_t gp = 0;
         for (uint32_t r = g.row0; r < g.row1; r++) gp += Q.rows[r].n_prm;
         // its primitive blocks and, for the matrix-core classes, its rows of
         // the contraction buffer (behind the chunk's blocks in the work buffer)
         gbytes[gi] = gp * sizeof(double) * (size_t)tp.ncp + (size_t)(g.row1 - g.row0) * ti.eco_row * sizeof(double);
         D.max_group_bytes = std::max(D.max_group_bytes, gbytes[gi]);
      }
      auto make_chunks = [&](size_t wbytes, std::vector<Chunk> &out, bool flops) {
         Chunk c{};
         size_t bytes = 0;
         for (size_t gi = 0; gi < D.n_groups; gi++) {
            const GroupRec &g = Q.groups[gi];
            const size_t gb = gbytes[gi];
            const size_t gp = (gb - (size_t)(g.row1 - g.row0) * ti.eco_row * sizeof(double)) /
                              (sizeof(double) * (size_t)tp.ncp);
            if (bytes + gb > wbytes && c.g1 > c.g0) {
               if (flops) D.max_chunk_bytes = std::max(D.max_chunk_bytes, bytes);
               out.push_back(c);
               c = Chunk{};
               c.g0 = c.g1 = (uint32_t)gi;
               c.r0 = c.r1 = g.row0;
               c.p0 = c.p1 = Q.rows[g.row0].prm_off;
               bytes = 0;
            }
            if (c.g1 == c.g0) {
               c.g0 = (uint32_t)gi;
               c.r0 = g.row0;
               c.p0 = Q.rows[g.row0].prm_off;
            }
            c.g1 = (uint32_t)gi + 1;
            c.r1 = g.row1;
            c.p1 = (uint32_t)(Q.rows[g.row0].prm_off + gp);
            bytes += gb;
            if (!flops) continue;
            // algorithmic FLOPs of the quartet kernel: contraction, HRR, SPH, KS
            for (int k = 0; k < tp.nclass; k++) {
               const double nq = ks_count(gks(gi, k).flags);
               const ClassPlan &cp = ti.cplans[k];
               const double contr = nq * 2.0 * (double)gp * cp.ncp;
               const double f = nq * ((double)(g.row1 - g.row0) * cp.hrr_flops + cp.sph_flops + 9.0 * cp.Ns);
               if (ti.lays[k].skip) {
                  c.tq_flops += contr + f;
               } else if (ti.lays[k].eco) {
                  c.eco_flops += contr;
                  c.k3_flops += f;
               } else {
                  c.k3_flops += contr + f;
               }
            }
         }
         if (c.g1 > c.g0) {
            if (flops) D.max_chunk_bytes = std::max(D.max_chunk_bytes, bytes);
            out.push_back(c);
         }
      };
      D.max_chunk_bytes = 0;
      make_chunks(work, D.chunks, true);
      D.chunks_na.clear();
      D.na_fit = D.max_group_bytes <= NOALLOC_AC0_BYTES;
      if (D.na_fit) make_chunks(NOALLOC_AC0_BYTES, D.chunks_na, false);
      n_rows_ += D.n_rows;
      if (release_host) Q.release();
   }
}

void Engine::ensure_work(size_t ncp_max, size_t max_prm_chunk, size_t max_chunk_bytes)
{
   // as large as the largest chunk (at most half the budget, or one set
   // quartet that is larger); grows, never shrinks (plan_evict frees it)
   const size_t want = std::max<size_t>(max_chunk_bytes, 1 << 20);
   if (!d_work_ || want > work_bytes_) {
      synchronize();
      if (d_work_) HIPCHECK(hipFree(d_work_));
      HIPCHECK(hipMalloc((void **)&d_work_, want));
      work_bytes_ = want;
   }
   if (!d_stats_) {
      HIPCHECK(hipMalloc((void **)&d_stats_, 3 * STATS_SLOTS * sizeof(unsigned long long)));
      HIPCHECK(hipMemset(d_stats_, 0, 3 * STATS_SLOTS * sizeof(unsigned long long)));
      HIPCHECK(hipMalloc((void **)&d_fixcnt_, 2 * sizeof(uint32_t)));
      HIPCHECK(hipMemset(d_fixcnt_, 0, 2 * sizeof(uint32_t)));
   }
   const size_t Ng = neighs_.size() / 3;
   const size_t cap = (debug_ & 1) ? std::max<size_t>(4096, max_prm_chunk * Ng) : 4096;
   if (cap > fix_cap_) {
      synchronize();
      if (d_fix_) HIPCHECK(hipFree(d_fix_));
      HIPCHECK(hipMalloc((void **)&d_fix_, 2 * sizeof(uint32_t) * cap));
      fix_cap_ = cap;
   }
   const size_t ws = fix_cap_ * (ncp_max + 16);
   if (ws > fix_ws_cap_) {
      synchronize();
      if (d_fix_ws_) HIPCHECK(hipFree(d_fix_ws_));
      HIPCHECK(hipMalloc((void **)&d_fix_ws_, sizeof(double) * ws));
      fix_ws_cap_ = ws;
   }
}

void Engine::ev_begin(int cat, void *stream)
{
   (void)cat;
   if (!timing_on_) return;
   ev_key_ = 0;
   if (ev_used_ + 2 > ev_pool_.size()) {
      hipEvent_t e0, e1;
      HIPCHECK(hipEventCreate(&e0));
      HIPCHECK(hipEventCreate(&e1));
      ev_pool_.push_back(e0);
      ev_pool_.push_back(e1);
   }
   ev_open_ = ev_pool_[ev_used_];
   HIPCHECK(hipEventRecord((hipEvent_t)ev_open_, (hipStream_t)stream));
}

void Engine::ev_end(int cat, void *stream)
{
   if (!timing_on_) return;
   void *e1 = ev_pool_[ev_used_ + 1];
   HIPCHECK(hipEventRecord((hipEvent_t)e1, (hipStream_t)stream));
   ev_live_.push_back(EvRec{cat, ev_open_, e1, ev_key_});
   ev_used_ += 2;
}

void Engine::harvest_events()
{
   for (const EvRec &r : ev_live_) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, (hipEvent_t)r.e0, (hipEvent_t)r.e1));
      timing_.ms[r.cat] += ms;
      timing_.launches[r.cat]++;
      detail_[r.cat][r.key & 255] += ms;
   }
   ev_live_.clear();
   ev_used_ = 0;
}

DevTables Engine::dev_tables() const
{
   DevTables t;
   t.C0 = d_C0_;
   t.ld_C0 = ld_C0_;
   t.grid_lo = d_grid_;
   t.grid_hi = d_grid_ + TCG_GRID * TCG_GRID;
   t.bw = d_bw_;
   t.ftable = d_ftab_;
   t.cell = d_cell_;
   t.neighs = d_neighs_;
   t.Ng = (int)neighs_.size() / 3;
   t.truncated = truncated_ ? 1 : 0;
   t.R_cut = R_cut_;
   return t;
}

void Engine::run_chunk(const TypeInfo &ti, const TypeDev &td, const Chunk &c, const RunBufs &w, const KTargets &kt)
{
   const TypePlan &tp = ti.plan;
   hipStream_t st = (hipStream_t)stream_;
   const DevTables t = dev_tables();
   BatchDev b{};   // every field zero / null unless set below
   b.tp = ti.d_plan;
   b.e1 = ti.d_e1;
   b.e2 = ti.d_e2;
   b.rows = td.rows;
   b.groups = td.groups;
   b.prm_row = td.prm_row;
   b.prm_ip = td.prm_ip;
   b.ksg = td.ksg;
   b.kss = td.kss;
   b.ksx = td.ksx;
   b.data = d_data_;
   b.ac0 = w.ac0;
   b.p0 = c.p0;
   b.p1 = c.p1;
   b.r0 = c.r0;
   b.r1 = c.r1;
   b.g0 = c.g0;
   b.g1 = c.g1;
   b.stats = w.stats;
   b.fix = w.fix;
   b.fix_count = w.fixcnt;
   b.fix_cap = w.fix_cap;
   b.force_fix = debug_ & 1;
   b.fix_ws = w.fix_ws;
   b.eco = w.ac0 + (size_t)(c.p1 - c.p0) * tp.ncp;
   b.eco_row = ti.eco_row;
   b.gmag = mixed_mode_ == 1 ? td.gmag : nullptr;
   b.gflag = mixed_mode_ == 2 ? td.gflag : nullptr;
   auto evb = [&](int cat, void *q) { if (w.count) ev_begin(cat, q); };
   auto eve = [&](int cat, void *q) { if (w.count) ev_end(cat, q); };

   if (launch_zero_u32(w.fixcnt, 1, st)) fatal("counter reset launch", 0);
   evb(0, st);
   ev_key_ = tp.L;
   int rc = launch_fm_r(tp, t, b, st);
   if (rc) fatal("fm_r launch", rc);
   eve(0, st);
   evb(3, st);
   rc = launch_fixup(tp, t, b, (debug_ & 1) ? w.fix_cap : std::min<uint32_t>(64u, w.fix_cap), st);
   if (rc) fatal("fix-up launch", rc);
   eve(3, st);
   if (mixed_mode_ == 2 && launch_mixed_flags(tp, b, td.gmag, d_mix_, mixed_tau_, td.gflag, d_mix_ + 1, st))
      fatal("mixed-precision route launch", 0);
   if (!ti.eco_items.empty()) {
      EcoArgs e{};
      e.items = ti.d_eco_items;
      e.n_items = (int)ti.eco_items.size();
      e.srcmaps = ti.d_srcmaps;
      e.src_off = ti.d_eco_info;
      e.eco_off = ti.d_eco_info + MAXTC;
      e.row_doubles = ti.eco_row;
      evb(4, st);
      ev_key_ = tp.L;
      rc = launch_eco(tp, b, e, st);
      if (rc) fatal("matrix-core contraction launch", rc);
      eve(4, st);
   }
   // debug bit 1: every consumer on the main stream, one after the other
   // (kernel durations without overlap, for profiling)
   const bool serial = (debug_ & 2) != 0;
   ensure_side();
   HIPCHECK(hipEventRecord((hipEvent_t)ev_join_[0], st));
   for (int i = 0; i < NSIDE; i++) HIPCHECK(hipStreamWaitEvent((hipStream_t)side_[i], (hipEvent_t)ev_join_[0], 0));
   for (int part = 0; part < GQS_PARTS; part++) {
      bool any = false;
      int npc = 0;
      for (int k = 0; k < tp.nclass; k++) {
         if (ti.lays[k].skip || ti.lays[k].part != part) continue;
         npc++;
         any = any || td.class_used[k];
      }
      if (!any) continue;
      void *q = (part == 0 || serial) ? (void *)st : side_[(part - 1) % NSIDE];
      evb(2, q);
      ev_key_ = tp.L;
      rc = launch_gqs(tp, ti.d_cps, ti.d_lays, ti.d_srcmaps, ti.cq_lds, ti.cq_gws, part, npc, t, b, kt, w.gws, q,
                      ti.cq_wide);
      if (rc) fatal("gqs launch", rc);
      eve(2, q);
   }
   int si = 0;
   for (int k = 0; k < tp.nclass; k++) {
      if (!td.class_used[k] || !ti.lays[k].skip) continue;
      void *q = serial ? (void *)st : side_[si++ % NSIDE];
      evb(1, q);
      ev_key_ = tp.L;
      rc = launch_tq(k, tp, t, b, kt, ti.d_srcmaps + ti.lays[k].o_src, q);
      if (rc) fatal("tq launch", rc);
      eve(1, q);
   }
   for (int i = 0; i < NSIDE; i++) {
      HIPCHECK(hipEventRecord((hipEvent_t)ev_join_[1 + i], (hipStream_t)side_[i]));
      HIPCHECK(hipStreamWaitEvent(st, (hipEvent_t)ev_join_[1 + i], 0));
   }
   if (!w.count) return;
   timing_.flops[1] += c.tq_flops;
   timing_.flops[2] += c.k3_flops;
   timing_.flops[4] += c.eco_flops;
   n_prmc_ += (uint64_t)(c.p1 - c.p0) * tp.nclass;
   n_slots_ += (uint64_t)(c.p1 - c.p0) * (uint64_t)t.Ng * tp.nclass;
}

// Device buffers and targets of an execution (allocating or growing what
// the work needs; nothing of this is captured in a graph); false when there
// is no work.
bool Engine::prepare_exec(std::vector<TypeDev> &devs, uint64_t max_index, RunBufs &w, KTargets &kt, uint64_t n)
{
   if (max_index >= n) fatal("add_qrtt offsets reach beyond the P / K size of set_P", (long long)max_index);
   ensure_tables();
   upload_data();
   size_t ncp_max = 0, max_prm = 0, scr = 0, max_gb = 0;
   for (auto &D : devs) {
      if (D.chunks.empty()) continue;
      const TypeInfo &ti = types_[D.type];
      if (truncated_ && ti.plan.L > nder_)
         fatal("the truncated-Coulomb table has fewer derivative orders than this quartet's total angular momentum",
               ti.plan.L);
      ncp_max = std::max(ncp_max, (size_t)ti.plan.ncp);
      max_gb = std::max(max_gb, D.max_chunk_bytes);
      for (auto &c : D.chunks) max_prm = std::max(max_prm, (size_t)(c.p1 - c.p0));
      scr = std::max(scr, (size_t)ti.cq_gws * SCRATCH_GRID * GQS_PARTS);
   }
   if (ncp_max == 0) return false;
   ensure_work(ncp_max, max_prm, max_gb);
   if (scr > scratch_cap_) {
      synchronize();
      if (d_scratch_) HIPCHECK(hipFree(d_scratch_));
      HIPCHECK(hipMalloc((void **)&d_scratch_, scr * sizeof(double)));
      scratch_cap_ = scr;
   }
   w.ac0 = d_work_;
   w.fix = d_fix_;
   w.fixcnt = d_fixcnt_;
   w.fix_ws = d_fix_ws_;
   w.fix_cap = (uint32_t)fix_cap_;
   w.stats = stats_on_ ? d_stats_ : nullptr;
   w.gws = Scratch{d_scratch_, scratch_cap_, SCRATCH_GRID};
   w.count = true;
   kt.nspin = nspin_;
   kt.hf_fac = hf_fac_;
   kt.out = d_out_;
   for (int s = 0; s < 2; s++) {
      kt.P[s] = d_P_[s];
      kt.K[s] = d_K_[s];
   }
   return true;
}

void Engine::launch_all(std::vector<TypeDev> &devs, const RunBufs &w, const KTargets &kt)
{
   if (mixed_mode_ == 1) {
      // calibration: every set quartet's magnitude from 0
      for (auto &D : devs)
         if (D.gmag && launch_zero_u32((uint32_t *)D.gmag, (int)(2 * D.n_groups), stream_))
            fatal("calibration reset launch", 0);
   } else if (mixed_mode_ == 2) {
      // max |P| of this density (both spins), and the routed count from 0
      if (launch_zero_u32((uint32_t *)(d_mix_ + 1), 2, stream_)) fatal("route counter reset", 0);
      for (int sp = 0; sp < kt.nspin; sp++)
         if (launch_absmax(kt.P[sp], FP_size_, d_mix_, sp == 0, stream_)) fatal("max |P| launch", 0);
   }
   // debug bit 2: wait for each quartet type and report it on stderr
   const bool progress = (debug_ & 4) != 0;
   const auto t0 = std::chrono::steady_clock::now();
   size_t nty = 0, done = 0;
   for (auto &D : devs) nty += D.chunks.empty() ? 0 : 1;
   for (auto &D : devs) {
      if (D.chunks.empty()) continue;
      for (auto &c : D.chunks) run_chunk(types_[D.type], D, c, w, kt);
      if (!progress) continue;
      synchronize();
      const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::fprintf(stderr, "libgint: quartet type %zu of %zu done (L %d, %zu set quartets, %zu chunks) at %.1f s\n",
                   ++done, nty, types_[D.type].plan.L, D.n_groups, D.chunks.size(), sec);
      std::fflush(stderr);
   }
}

void Engine::execute(std::vector<TypeDev> &devs, uint64_t max_index)
{
   RunBufs w{};
   KTargets kt{};
   if (!prepare_exec(devs, max_index, w, kt, FP_size_)) return;
   if (debug_ & 24) force_kron(devs, (debug_ & 16) != 0);
   launch_all(devs, w, kt);
}

// ---------------------------------------------------------------------------
// the resident plan's launches as one HIP graph, captured at its first
// execution and replayed while the buffers and targets it was captured with
// are unchanged
// ---------------------------------------------------------------------------
void Engine::drop_graph()
{
   if (graph_exec_) {
      if (stream_) (void)hipStreamSynchronize((hipStream_t)stream_);
      (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
   }
   graph_exec_ = nullptr;
}

void Engine::replay_graph(const RunBufs &w, const KTargets &kt)
{
   const uint64_t key[] = {(uint64_t)w.ac0, (uint64_t)w.fix, (uint64_t)w.fix_ws, (uint64_t)w.fixcnt,
                           (uint64_t)w.gws.ptr, (uint64_t)w.gws.doubles, (uint64_t)w.fix_cap,
                           (uint64_t)kt.P[0], (uint64_t)kt.P[1], (uint64_t)kt.K[0], (uint64_t)kt.K[1],
                           (uint64_t)kt.nspin, graph_bits(kt.hf_fac), (uint64_t)kt.out, (uint64_t)d_data_,
                           (uint64_t)d_C0_, (uint64_t)d_neighs_, (uint64_t)neighs_.size(), graph_bits(R_cut_),
                           (uint64_t)truncated_, (uint64_t)debug_, (uint64_t)ld_C0_, (uint64_t)mixed_mode_,
                           graph_bits(mixed_tau_), (uint64_t)d_mix_};
   constexpr size_t NK = sizeof key / sizeof key[0];
   static_assert(NK <= sizeof graph_key_ / sizeof graph_key_[0], "graph key");
   bool same = graph_exec_ != nullptr;
   for (size_t i = 0; same && i < NK; i++) same = graph_key_[i] == key[i];
   if (!same) {
      drop_graph();
      const uint64_t s0 = n_slots_, p0 = n_prmc_;
      const double f1 = timing_.flops[1], f2 = timing_.flops[2], f4 = timing_.flops[4];
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture((hipStream_t)stream_, hipStreamCaptureModeThreadLocal));
      launch_all(plan_dev_, w, kt);
      HIPCHECK(hipStreamEndCapture((hipStream_t)stream_, &g));
      hipGraphExec_t ex = nullptr;
      HIPCHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      HIPCHECK(hipGraphDestroy(g));
      graph_exec_ = ex;
      for (size_t i = 0; i < NK; i++) graph_key_[i] = key[i];
      // host-side counts of one execution, added at every replay
      graph_counts_[0] = n_slots_ - s0;
      graph_counts_[1] = n_prmc_ - p0;
      graph_flops_[0] = timing_.flops[1] - f1;
      graph_flops_[1] = timing_.flops[2] - f2;
      graph_flops_[2] = timing_.flops[4] - f4;
      n_slots_ = s0;
      n_prmc_ = p0;
      timing_.flops[1] = f1;
      timing_.flops[2] = f2;
      timing_.flops[4] = f4;
      graph_captures_++;
   }
   HIPCHECK(hipGraphLaunch((hipGraphExec_t)graph_exec_, (hipStream_t)stream_));
   n_slots_ += graph_counts_[0];
   n_prmc_ += graph_counts_[1];
   timing_.flops[1] += graph_flops_[0];
   timing_.flops[2] += graph_flops_[1];
   timing_.flops[4] += graph_flops_[2];
   graph_replays_++;
}

// The resident plan, like plan_execute, but into the caller's device P and
// K (K is added to, not zeroed) and with every scratch buffer carved from
// the static no-allocation array: no device allocation, no memset node.
int Engine::plan_execute_noalloc(const double *Pa, const double *Pb, double *Ka, double *Kb, uint64_t n)
{
   auto err = [](int code, const char *what, long long v) {
      std::fprintf(stderr, "libgint: plan_execute_noalloc: %s (%lld)\n", what, v);
      std::fflush(stderr);
      return code;
   };
   if (recording_) return err(-1, "called while recording a plan", 0);
   if (!stream_) return err(-1, "libgint_init has not been called", 0);
   if (!Pa || !Ka || (!Pb) != (!Kb)) return err(-2, "P and K device arrays are required (both spins or neither)", 0);
   if (cur_type_ >= 0 || !touched_.empty()) return err(-1, "a set quartet is open: call add_set first", 0);
   bool any = false;
   size_t ncp_max = 0;
   for (auto &D : plan_dev_) {
      if (D.chunks.empty()) continue;   // a type with no recorded work
      any = true;
      const TypeInfo &ti = types_[D.type];
      if (!D.na_fit || D.chunks_na.empty())
         return err(-4, "a set quartet's primitive blocks exceed the no-allocation scratch (bytes of the largest)",
                    (long long)D.max_group_bytes);
      if (truncated_ && ti.plan.L > nder_) return err(-7, "the C0 table has too few derivative orders", ti.plan.L);
      if (ti.cq_gws && NOALLOC_GWS_BYTES / 8 / ((size_t)GQS_PARTS * ti.cq_gws) < 1)
         return err(-5, "a class workspace exceeds the no-allocation scratch", ti.cq_gws);
      ncp_max = std::max(ncp_max, (size_t)ti.plan.ncp);
   }
   if (!any) return 0;
   if (plan_max_index_ >= n) return err(-3, "the plan's add_qrtt offsets reach beyond n", (long long)plan_max_index_);
   if (tables_dirty_ || !d_grid_ || data_version_dev_ != store_.version())
      return err(-6, "tables or basis data changed since libgint_amd_plan_end: record the plan again", 0);
   if (ncp_max > (size_t)MAX_NCP) return err(-4, "block size beyond the fix-up workspace", (long long)ncp_max);
   char *base = (char *)noalloc_scratch_;
   if (!base) return err(-8, "the no-allocation scratch array has no device address", 0);
   RunBufs w{};
   w.ac0 = (double *)(base + NOALLOC_OFF_AC0);
   w.fix = (uint32_t *)(base + NOALLOC_OFF_FIX);
   w.fixcnt = (uint32_t *)base;
   w.fix_ws = (double *)(base + NOALLOC_OFF_FIXWS);
   w.fix_cap = (uint32_t)NOALLOC_FIX;
   w.stats = nullptr;
   w.count = false;
   KTargets kt{};
   kt.nspin = Pb ? 2 : 1;
   kt.hf_fac = hf_fac_;
   kt.out = d_out_;
   kt.P[0] = Pa;
   kt.P[1] = Pb;
   kt.K[0] = Ka;
   kt.K[1] = Kb;
   // the overflow word (it is only read back by the allocating path)
   if (launch_zero_u32(w.fixcnt, 2, stream_)) fatal("counter reset launch", 0);
   for (auto &D : plan_dev_) {
      if (D.chunks.empty()) continue;
      const TypeInfo &ti = types_[D.type];
      const unsigned g = ti.cq_gws ? (unsigned)std::min<size_t>(SCRATCH_GRID, NOALLOC_GWS_BYTES / 8 /
                                                                                ((size_t)GQS_PARTS * ti.cq_gws))
                                   : SCRATCH_GRID;
      w.gws = Scratch{(double *)(base + NOALLOC_OFF_GWS), NOALLOC_GWS_BYTES / 8, g};
      for (auto &c : D.chunks_na) run_chunk(ti, D, c, w, kt);
   }
   return 0;
}

// ---------------------------------------------------------------------------
// autotuner: per class of a quartet type, the schedule of its contraction,
// HRR, spherical transform and K update, chosen by timing on a sample of the
// plan (first execution after plan_end):
//   0  the templated one-wavefront-per-set-quartet kernel (k_tqw), where
//      the class has one,
//   1  the workgroup kernel (k_gqs) with its own contraction loop,
//   2  k_gqs after the matrix-core contraction (k_eco), where the class has
//      a place in the contraction buffer.
// Decisions are kept per (class, contractions, primitive quartets per row
// in powers of two) for the life of the process and reused by later plans.
// ---------------------------------------------------------------------------
static std::mutex g_tune_mu;
static std::map<uint64_t, int> g_tune;   // key -> schedule

static uint64_t tune_key(int code, int nq, size_t n_prm, size_t n_rows)
{
   const double avg = n_rows ? (double)n_prm / (double)n_rows : 1.0;
   const int bucket = std::min(15, std::max(0, (int)std::floor(std::log2(std::max(avg, 1.0)))));
   return (uint64_t)code | (uint64_t)nq << 8 | (uint64_t)bucket << 20;
}

void Engine::apply_schedule(TypeInfo &ti, int k, int v)
{
   CqLayout &y = ti.lays[k];
   y.skip = v == 0 ? 1 : 0;
   y.eco = (v == 2 || v == 3) ? 1 : 0;
   y.kron = v >= 3 ? 1 : 0;
}

// debug bits 3 and 4: every class (of a plan, or of the call stream's
// types) that can take the Kronecker path takes it, whatever the schedule: bit 3 with the matrix-core
// contraction where the class has a place in the contraction buffer (else
// its own), bit 4 with its own contraction where that fits (else the matrix
// cores)
void Engine::force_kron(std::vector<TypeDev> &devs, bool own_first)
{
   for (auto &D : devs) {
      if (D.chunks.empty()) continue;
      TypeInfo &ti = types_[D.type];
      bool any = false;
      for (CqLayout &y : ti.lays) {
         if (!y.kron_ws) continue;
         const bool can_eco = y.eco_off >= 0, can_own = y.o_kc >= 0;
         if (!can_eco && !can_own) continue;
         const int eco = (can_eco && (!own_first || !can_own)) ? 1 : 0;
         if (y.kron && !y.skip && y.eco == eco) continue;
         y.skip = 0;
         y.eco = eco;
         y.kron = 1;
         any = true;
      }
      if (any) upload_layouts(ti);
   }
}

void Engine::upload_layouts(TypeInfo &ti)
{
   ti.cq_wide = false;
   for (const CqLayout &y : ti.lays) ti.cq_wide = ti.cq_wide || (y.kron && !y.skip);
   assign_parts(ti.plan, ti.cplans, ti.lays);
   make_eco_items(ti);
   HIPCHECK(hipMemcpy(ti.d_lays, ti.lays.data(), sizeof(CqLayout) * ti.lays.size(), hipMemcpyHostToDevice));
   if (ti.d_eco_items) HIPCHECK(hipFree(ti.d_eco_items));
   ti.d_eco_items = nullptr;
   if (!ti.eco_items.empty()) {
      HIPCHECK(hipMalloc((void **)&ti.d_eco_items, sizeof(EcoItem) * ti.eco_items.size()));
      HIPCHECK(hipMemcpy(ti.d_eco_items, ti.eco_items.data(), sizeof(EcoItem) * ti.eco_items.size(),
                         hipMemcpyHostToDevice));
   }
}

void Engine::autotune(const RunBufs &w, const KTargets &kt)
{
   synchronize();
   const auto t0 = std::chrono::steady_clock::now();
   // a scratch K: the timed launches must not touch the caller's
   double *Ks = nullptr;
   HIPCHECK(hipMalloc((void **)&Ks, sizeof(double) * std::max<size_t>(FP_size_, 1) * (size_t)kt.nspin));
   KTargets ks = kt;
   ks.K[0] = Ks;
   ks.K[1] = kt.nspin > 1 ? Ks + FP_size_ : nullptr;
   ks.out = nullptr;
   hipStream_t st = (hipStream_t)stream_;
   hipEvent_t e0, e1;
   HIPCHECK(hipEventCreate(&e0));
   HIPCHECK(hipEventCreate(&e1));
   EcoItem *d_items = nullptr;
   size_t items_cap = 0;
   const DevTables t = dev_tables();
   for (auto &D : plan_dev_) {
      if (D.chunks.empty()) continue;
      TypeInfo &ti = types_[D.type];
      const TypePlan &tp = ti.plan;
      // decisions already known for every class of the type
      std::vector<int> dec(tp.nclass, -1);
      bool all = true;
      {
         std::lock_guard<std::mutex> lk(g_tune_mu);
         for (int k = 0; k < tp.nclass; k++) {
            auto it = g_tune.find(tune_key(tp.cls_code[k], tp.cls_nl[k], D.n_prm, D.n_rows));
            if (it != g_tune.end()) dec[k] = it->second;
            else if (D.class_used[k]) all = false;
         }
      }
      if (!all) {
         // the sample: the first set quartets of the first chunk, about
         // TUNE_PRM primitive quartets (launches long enough to time), at
         // most TUNE_GROUPS set quartets
         const Chunk &c0 = D.chunks[0];
         Chunk c = c0;
         const double ppg = (double)(c0.p1 - c0.p0) / (double)std::max<uint32_t>(c0.g1 - c0.g0, 1);
         const uint32_t ng = (uint32_t)std::max(1.0, std::min<double>(TUNE_GROUPS, TUNE_PRM / std::max(ppg, 1.0)));
         c.g1 = std::min<uint32_t>(c0.g1, c0.g0 + ng);
         if (c.g1 < c0.g1) {
            GroupRec gl;
            HIPCHECK(hipMemcpy(&gl, D.groups + (c.g1 - 1), sizeof gl, hipMemcpyDeviceToHost));
            RowRec rl;
            HIPCHECK(hipMemcpy(&rl, D.rows + (gl.row1 - 1), sizeof rl, hipMemcpyDeviceToHost));
            c.r1 = gl.row1;
            c.p1 = rl.prm_off + rl.n_prm;
         }
         BatchDev b{};   // every field zero / null unless set below
         b.tp = ti.d_plan;
         b.e1 = ti.d_e1;
         b.e2 = ti.d_e2;
         b.rows = D.rows;
         b.groups = D.groups;
         b.prm_row = D.prm_row;
         b.prm_ip = D.prm_ip;
         b.ksg = D.ksg;
         b.kss = D.kss;
         b.ksx = D.ksx;
         b.data = d_data_;
         b.ac0 = w.ac0;
         b.p0 = c.p0;
         b.p1 = c.p1;
         b.r0 = c.r0;
         b.r1 = c.r1;
         b.g0 = c.g0;
         b.g1 = c.g1;
         b.stats = nullptr;
         b.fix = w.fix;
         b.fix_count = w.fixcnt;
         b.fix_cap = w.fix_cap;
         b.force_fix = 0;
         b.fix_ws = w.fix_ws;
         b.eco = w.ac0 + (size_t)(c.p1 - c.p0) * tp.ncp;
         b.eco_row = ti.eco_row;
         b.gmag = nullptr;    // no calibration and no single-precision route while timing
         b.gflag = nullptr;
         if (launch_zero_u32(w.fixcnt, 1, st) || launch_fm_r(tp, t, b, st) ||
             launch_fixup(tp, t, b, std::min<uint32_t>(64u, w.fix_cap), st))
            fatal("autotune: sample launch", 0);
         const std::vector<CqLayout> keep = ti.lays;
         for (int k = 0; k < tp.nclass; k++) {
            if (!D.class_used[k] || dec[k] >= 0) continue;
            double best = 1e300;
            int bv = keep[k].skip ? 0 : (keep[k].eco ? (keep[k].kron ? 3 : 2) : (keep[k].kron ? 4 : 1));
            // the sample's contraction FLOPs of the class: above TUNE_VALU_FLOPS
            // the VALU contraction loop (schedule 1) is not timed where the
            // matrix cores can take the class (it is slower there by an order
            // of magnitude or two, and timing it would cost seconds)
            const double cflops = 2.0 * tp.cls_nl[k] * (double)tp.cls_ncp[k] * (double)(c.p1 - c.p0);
            for (int v = 0; v < 5; v++) {
               if (v == 0 && !tq_has(tp.cls_code[k])) continue;
               if ((v == 2 || v == 3) && keep[k].eco_off < 0) continue;
               if (v >= 3 && !keep[k].kron_ws) continue;
               if (v == 4 && keep[k].o_kc < 0) continue;
               // the own contraction is not timed where the VALU
               // contraction would be too slow (as schedule 1)
               if (v == 4 && keep[k].eco_off >= 0 && cflops > TUNE_VALU_FLOPS) continue;
               if (v == 1 && keep[k].eco_off >= 0 && cflops > TUNE_VALU_FLOPS) continue;
               // this class alone in part 0 with schedule v
               std::vector<CqLayout> L = keep;
               for (auto &y : L) y.part = -1;
               L[k].skip = v == 0 ? 1 : 0;
               L[k].eco = (v == 2 || v == 3) ? 1 : 0;
               L[k].kron = v >= 3 ? 1 : 0;
               L[k].part = 0;
               HIPCHECK(hipMemcpy(ti.d_lays, L.data(), sizeof(CqLayout) * L.size(), hipMemcpyHostToDevice));
               std::vector<EcoItem> items;
               if (v == 2 || v == 3) {
                  for (int jt0 = 0; jt0 * 16 < tp.cls_ncp[k]; jt0 += 4)
                     for (int qb = 0; qb * 64 < tp.cls_nl[k]; qb++)
                        items.push_back(EcoItem{(uint16_t)k, (uint16_t)jt0, (uint16_t)qb, 0});
                  if (items.size() > items_cap) {
                     if (d_items) HIPCHECK(hipFree(d_items));
                     items_cap = items.size() * 2;
                     HIPCHECK(hipMalloc((void **)&d_items, sizeof(EcoItem) * items_cap));
                  }
                  HIPCHECK(hipMemcpy(d_items, items.data(), sizeof(EcoItem) * items.size(), hipMemcpyHostToDevice));
               }
               // the faster of two runs: the first may load the kernel's code object
               float ms = 1e30f;
               for (int rep = 0; rep < 2; rep++) {
                  float msr = 0.f;
                  HIPCHECK(hipEventRecord(e0, st));
                  int rc = 0;
                  if (v == 0) {
                     rc = launch_tq(k, tp, t, b, ks, ti.d_srcmaps + keep[k].o_src, st);
                  } else {
                     if (v == 2 || v == 3) {
                        EcoArgs e{};
                        e.items = d_items;
                        e.n_items = (int)items.size();
                        e.srcmaps = ti.d_srcmaps;
                        e.src_off = ti.d_eco_info;
                        e.eco_off = ti.d_eco_info + MAXTC;
                        e.row_doubles = ti.eco_row;
                        rc = launch_eco(tp, b, e, st);
                     }
                     if (!rc)
                        rc = launch_gqs(tp, ti.d_cps, ti.d_lays, ti.d_srcmaps, ti.cq_lds, ti.cq_gws, 0, 1, t, b, ks,
                                        w.gws, st, v >= 3 || ti.cq_wide);
                  }
                  if (rc) fatal("autotune: candidate launch", rc);
                  HIPCHECK(hipEventRecord(e1, st));
                  HIPCHECK(hipEventSynchronize(e1));
                  HIPCHECK(hipEventElapsedTime(&msr, e0, e1));
                  ms = std::min(ms, msr);
               }
               if (ms < best) {
                  best = ms;
                  bv = v;
               }
            }
            dec[k] = bv;
            tune_ms_[tp.cls_code[k]] += best;
            tune_pick_[tp.cls_code[k]] = bv;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[tune_key(tp.cls_code[k], tp.cls_nl[k], D.n_prm, D.n_rows)] = bv;
         }
      }
      for (int k = 0; k < tp.nclass; k++)
         if (dec[k] >= 0) {
            apply_schedule(ti, k, dec[k]);
            tune_counts_[dec[k]]++;
         }
      upload_layouts(ti);
      if (debug_ & 4) {
         std::fprintf(stderr, "libgint: autotune: quartet type L %d (%zu set quartets) %s at %.1f s\n", tp.L, D.n_groups,
                      all ? "from earlier decisions" : "timed",
                      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
         std::fflush(stderr);
      }
   }
   HIPCHECK(hipStreamSynchronize(st));
   if (d_items) HIPCHECK(hipFree(d_items));
   HIPCHECK(hipFree(Ks));
   HIPCHECK(hipEventDestroy(e0));
   HIPCHECK(hipEventDestroy(e1));
   tune_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   plan_tuned_ = true;
}

// ---------------------------------------------------------------------------
// serving (server.cpp): a client handle with a stream of its own, whose
// resident plan can leave the device (host copy) and come back, executed
// into device P / K arrays that the server keeps per client
// ---------------------------------------------------------------------------
void Engine::use_private_stream()
{
   if (own_stream_) return;
   synchronize();
   hipStream_t q = nullptr;
   HIPCHECK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
   stream_ = q;
   own_stream_ = true;
   ensure_side();
}

uint64_t Engine::device_bytes() const
{
   uint64_t b = 0;
   for (const auto &d : plan_dev_)
      b += d.cap[0] * sizeof(RowRec) + d.cap[1] * sizeof(GroupRec) + d.cap[2] * 4 + d.cap[3] * 4 +
           d.cap[4] * sizeof(KsGroup) + d.cap[5] * sizeof(KsRec) + d.cap[6] * sizeof(KsSet);
   b += work_bytes_ + fix_cap_ * 8 + fix_ws_cap_ * 8 + scratch_cap_ * 8 + d_data_cap_ * 8 + C0_.size() * 8 +
        d_neighs_cap_ * 8;
   return b;
}

uint64_t Engine::work_estimate() const
{
   size_t ncp_max = 0, mcb = 0, scr = 0;
   for (const auto &D : plan_dev_) {
      if (D.chunks.empty()) continue;
      const TypeInfo &ti = types_[D.type];
      ncp_max = std::max(ncp_max, (size_t)ti.plan.ncp);
      mcb = std::max(mcb, D.max_chunk_bytes);
      scr = std::max(scr, (size_t)ti.cq_gws * SCRATCH_GRID * GQS_PARTS);
   }
   if (!ncp_max) return 0;
   return std::max<size_t>(mcb, 1 << 20) + 2 * 4 * 4096 + 4096 * (ncp_max + 16) * 8 + scr * 8;
}

uint64_t Engine::plan_host_bytes() const
{
   uint64_t b = 0;
   for (const auto &Q : plan_host_)
      b += Q.rows.size() * sizeof(RowRec) + Q.groups.size() * sizeof(GroupRec) + Q.prm_ip.size() * 8 +
           Q.ksg.size() * sizeof(KsGroup) + Q.ksx.size() * sizeof(KsRec) + Q.kss.size() * sizeof(KsSet);
   return b;
}

template <class T>
static void dev_get(std::vector<T> &v, const T *ptr, size_t n)
{
   v.resize(n);
   if (n) HIPCHECK(hipMemcpy(v.data(), ptr, sizeof(T) * n, hipMemcpyDeviceToHost));
}

void Engine::plan_evict()
{
   if (recording_) fatal("plan_evict while recording", 0);
   synchronize();
   drop_graph();
   plan_host_.assign(plan_dev_.size(), TypeQueue());
   for (size_t ty = 0; ty < plan_dev_.size(); ty++) {
      TypeDev &D = plan_dev_[ty];
      if (D.n_groups) {
         TypeQueue &Q = plan_host_[ty];
         dev_get(Q.rows, D.rows, D.n_rows);
         dev_get(Q.groups, D.groups, D.n_groups);
         dev_get(Q.prm_ip, D.prm_ip, D.n_prm);
         dev_get(Q.prm_row, D.prm_row, D.n_prm);
         dev_get(Q.ksg, D.ksg, D.n_ksg);
         dev_get(Q.ksx, D.ksx, D.n_ksx);
         dev_get(Q.kss, D.kss, D.n_kss);
      }
      D.release();
   }
   plan_dev_.clear();
   mixed_calibrated_ = false;
   void *bufs[] = {d_work_, d_fix_, d_fix_ws_, d_scratch_};
   for (void *b : bufs)
      if (b) HIPCHECK(hipFree(b));
   d_work_ = nullptr;
   d_fix_ = nullptr;
   d_fix_ws_ = nullptr;
   d_scratch_ = nullptr;
   work_bytes_ = fix_cap_ = fix_ws_cap_ = scratch_cap_ = 0;
   evicted_ = true;
}

void Engine::plan_restore()
{
   if (!evicted_) return;
   upload(plan_host_, plan_dev_, true);
   plan_host_.clear();
   evicted_ = false;
}

void Engine::plan_execute_into(const double *P0, const double *P1, double *K0, double *K1, int nspin, uint64_t n)
{
   if (recording_) fatal("plan_execute_into while recording", 0);
   if (evicted_) fatal("plan_execute_into with the plan off the device", 0);
   RunBufs w{};
   KTargets kt{};
   if (!prepare_exec(plan_dev_, plan_max_index_, w, kt, n)) return;
   kt.nspin = nspin;
   kt.P[0] = P0;
   kt.P[1] = P1;
   kt.K[0] = K0;
   kt.K[1] = K1;
   if (graph_on_ && nomp_ == 1 && !timing_on_ && !stats_on_ && (debug_ & 6) == 0)
      replay_graph(w, kt);
   else
      launch_all(plan_dev_, w, kt);
}

// ---------------------------------------------------------------------------
// counters
// ---------------------------------------------------------------------------
void Engine::reset_counters()
{
   synchronize();
   n_slots_ = n_prmc_ = n_q_ = n_rows_ = 0;
   timing_ = Timing();
   if (d_stats_) HIPCHECK(hipMemset(d_stats_, 0, 3 * STATS_SLOTS * sizeof(unsigned long long)));
   for (auto &d : detail_) d.fill(0.0);
}

Timing Engine::timing()
{
   synchronize();
   return timing_;
}

void Engine::timing_detail(int cat, double out[256])
{
   synchronize();
   for (int i = 0; i < 256; i++) out[i] = (cat >= 0 && cat < 5) ? detail_[cat][i] : 0.0;
}

void Engine::stats(uint64_t out[8])
{
   synchronize();
   unsigned long long s[3] = {0, 0, 0};
   if (d_stats_) {
      std::vector<unsigned long long> part(3 * STATS_SLOTS);
      HIPCHECK(hipMemcpy(part.data(), d_stats_, sizeof(unsigned long long) * part.size(), hipMemcpyDeviceToHost));
      for (unsigned i = 0; i < STATS_SLOTS; i++)
         for (int k = 0; k < 3; k++) s[k] += part[3 * i + k];
   }
   out[0] = n_slots_;
   out[1] = s[0];
   out[2] = n_prmc_;
   out[3] = n_q_;
   out[4] = s[1];
   out[5] = n_rows_;
   out[6] = s[2];
   out[7] = 0;
}

}  // namespace gint
