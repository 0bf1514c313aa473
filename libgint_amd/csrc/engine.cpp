// Host engine behind the libgint_* C-ABI.  See engine.hpp.
//
// Built with g++ -fopenmp (libgomp) so that the OpenMP-collective calls of
// the protocol (init, set_P*, set_K*, get_K*) synchronise with a caller
// built by gfortran/gcc, like CP2K.
#include "engine.hpp"

#include <hip/hip_runtime_api.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gint_math.h"

namespace gint {

void fatal(const char *what, int code)
{
   std::fprintf(stderr, "libgint: %s (code %d)\n", what, code);
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
   if (d_data_) (void)hipFree(d_data_);
   if (d_C0_) (void)hipFree(d_C0_);
   if (d_grid_) (void)hipFree(d_grid_);
   if (d_bw_) (void)hipFree(d_bw_);
   if (d_ftab_) (void)hipFree(d_ftab_);
   if (d_cell_) (void)hipFree(d_cell_);
   if (d_neighs_) (void)hipFree(d_neighs_);
   if (d_arena_) (void)hipFree(d_arena_);
   if (h_stage_) (void)hipHostFree(h_stage_);
   if (d_scratch_) (void)hipFree(d_scratch_);
   if (d_stats_) (void)hipFree(d_stats_);
   for (auto &e : ev_)
      if (e) (void)hipEventDestroy((hipEvent_t)e);
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
}

void Engine::set_potential_truncated(double R_cut, const double *C0, int ld_C0, int C0_size)
{
   if (ld_C0 % TCG_NCOEF != 0 || C0_size < ld_C0 * TCG_NPATCH) fatal("set_Potential_Truncated: bad C0 table shape", ld_C0);
   truncated_ = true;
   R_cut_ = R_cut;
   ld_C0_ = ld_C0;
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
   if (i < 0) fatal("set_Atom: negative set index", i);
   if (np < 1 || np > MAXNP) fatal("set_Atom: 1..16 primitives per set", np);
   if (i >= (int)sets_.size()) sets_.resize(i + 1);
   SetInfo &s = sets_[i];
   s.R = store_.add(R, 3);
   s.Z = store_.add(Z, np);
   s.np = np;
   s.l.clear();
   s.nl.clear();
   s.K.clear();
}

void Engine::set_atom_L(int i, int l, int nl, const double *K)
{
   if (i < 0 || i >= (int)sets_.size() || sets_[i].np == 0) fatal("set_Atom_L before set_Atom", i);
   if (l < 0 || l > MAXL) fatal("set_Atom_L: l must be 0..3", l);
   if (nl < 1 || nl > MAXNL) fatal("set_Atom_L: 1..4 contractions per shell", nl);
   SetInfo &s = sets_[i];
   s.l.push_back(l);
   s.nl.push_back(nl);
   s.K.push_back(store_.add(K, nl * s.np));
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
   if (stream_) HIPCHECK(hipStreamSynchronize((hipStream_t)stream_));
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
   if (prm_tmp_.empty()) return;
   const int ns = (int)sets_.size();
   if (i < 0 || j < 0 || k < 0 || l < 0 || i >= ns || j >= ns || k >= ns || l >= ns) fatal("add_shell: unknown set", i);
   const int Ng = (int)neighs_.size() / 3;
   if (n1 < 0 || n2 < 0 || n1 >= Ng || n2 >= Ng) fatal("add_shell: cell index beyond set_neighs", n1 > n2 ? n1 : n2);
   const SetInfo *S[4] = {&sets_[i], &sets_[j], &sets_[k], &sets_[l]};
   for (int t = 0; t < 4; t++)
      for (uint32_t ip : prm_tmp_) {
         const int v = (ip >> (4 * t)) & 15;
         if (v >= S[t]->np) fatal("add_prm: primitive index beyond the set's primitives", v);
      }
   const uint32_t np_bits = (uint32_t)(S[0]->np - 1) << 16 | (uint32_t)(S[1]->np - 1) << 20 |
                            (uint32_t)(S[2]->np - 1) << 24 | (uint32_t)(S[3]->np - 1) << 28;
   for (size_t sa = 0; sa < S[0]->l.size(); sa++)
      for (size_t sb = 0; sb < S[1]->l.size(); sb++)
         for (size_t sc = 0; sc < S[2]->l.size(); sc++)
            for (size_t sd = 0; sd < S[3]->l.size(); sd++) {
               const int c = class_index(S[0]->l[sa], S[1]->l[sb], S[2]->l[sc], S[3]->l[sd]);
               ClassQueue &Q = q_[c];
               if (!Q.touched) {
                  Q.touched = true;
                  Q.set_row0 = Q.rows.size();
                  Q.set_prm0 = Q.prm_ip.size();
                  Q.set_ks0 = Q.ks.size();
                  Q.set_nq = 0;
                  Q.set_eslots0 = Q.e_slots;
                  set_classes_.push_back(c);
               }
               const int nla = S[0]->nl[sa], nlb = S[1]->nl[sb], nlc = S[2]->nl[sc], nld = S[3]->nl[sd];
               RowRec r;
               r.prm_off = (uint32_t)Q.prm_ip.size();
               r.n_prm = (uint32_t)prm_tmp_.size();
               r.q_off = Q.n_q + Q.set_nq;
               r.nlnp = (uint32_t)nla | (uint32_t)nlb << 4 | (uint32_t)nlc << 8 | (uint32_t)nld << 12 | np_bits;
               r.n12 = (uint32_t)n1 | (uint32_t)n2 << 16;
               for (int t = 0; t < 4; t++) {
                  r.R[t] = S[t]->R;
                  r.Z[t] = S[t]->Z;
               }
               r.K[0] = S[0]->K[sa];
               r.K[1] = S[1]->K[sb];
               r.K[2] = S[2]->K[sc];
               r.K[3] = S[3]->K[sd];
               r.e_off = Q.e_slots;
               r.set_ij = (uint32_t)(i & 0xffff) | (uint32_t)(j & 0xffff) << 16;
               r.set_kl = (uint32_t)(k & 0xffff) | (uint32_t)(l & 0xffff) << 16;
               Q.e_slots += (uint32_t)(nla * nlb * nlc * nld);
               const uint32_t row_idx = (uint32_t)Q.rows.size();
               Q.rows.push_back(r);
               Q.prm_ip.insert(Q.prm_ip.end(), prm_tmp_.begin(), prm_tmp_.end());
               Q.prm_row.insert(Q.prm_row.end(), prm_tmp_.size(), row_idx);
            }
   prm_tmp_.clear();
}

void Engine::add_qrt(int la, int lb, int lc, int ld, int nla, int nlb, int nlc, int nld)
{
   if (la < 0 || la > MAXL || lb < 0 || lb > MAXL || lc < 0 || lc > MAXL || ld < 0 || ld > MAXL)
      fatal("add_qrt: l must be 0..3", la);
   const int c = class_index(la, lb, lc, ld);
   ClassQueue &Q = q_[c];
   if (!Q.touched) {
      Q.touched = true;
      Q.set_row0 = Q.rows.size();
      Q.set_prm0 = Q.prm_ip.size();
      Q.set_ks0 = Q.ks.size();
      Q.set_nq = 0;
      Q.set_eslots0 = Q.e_slots;
      set_classes_.push_back(c);
   }
   Q.set_nq += (uint32_t)(nla * nlb * nlc * nld);
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
   (void)inla; (void)inlb; (void)inlc; (void)inld;
   if (la < 0 || la > MAXL || lb < 0 || lb > MAXL || lc < 0 || lc > MAXL || ld < 0 || ld > MAXL)
      fatal("add_qrtt: l must be 0..3", la);
   if (ld_ac < 0 || ld_ad < 0 || ld_bc < 0 || ld_bd < 0 || ld_ac > 0xffff || ld_ad > 0xffff || ld_bc > 0xffff ||
       ld_bd > 0xffff)
      fatal("add_qrtt: leading dimension out of 0..65535", ld_ac);
   const int c = class_index(la, lb, lc, ld);
   ClassQueue &Q = q_[c];
   if (!Q.touched) {
      Q.touched = true;
      Q.set_row0 = Q.rows.size();
      Q.set_prm0 = Q.prm_ip.size();
      Q.set_ks0 = Q.ks.size();
      Q.set_nq = 0;
      Q.set_eslots0 = Q.e_slots;
      set_classes_.push_back(c);
   }
   KsRec e;
   e.off[0] = off_ac;
   e.off[1] = off_ad;
   e.off[2] = off_bc;
   e.off[3] = off_bd;
   e.ld01 = (uint32_t)ld_ac | (uint32_t)ld_ad << 16;
   e.ld23 = (uint32_t)ld_bc | (uint32_t)ld_bd << 16;
   e.flags = (uint32_t)Tac | (uint32_t)Tad << 1 | (uint32_t)Tbc << 2 | (uint32_t)Tbd << 3 | symm_code(symm_fac) << 4;
   e.row0 = e.row1 = 0;
   e.l = 0;
   e.pad[0] = e.pad[1] = 0;
   Q.ks.push_back(e);
}

void Engine::add_set()
{
   for (int c : set_classes_) {
      ClassQueue &Q = q_[c];
      const uint32_t base = Q.n_q;
      // attach each quartet of the set to the rows that produce it
      const size_t r0 = Q.set_row0, r1 = Q.rows.size();
      for (size_t e = Q.set_ks0; e < Q.ks.size(); e++) {
         KsRec &k = Q.ks[e];
         const uint32_t q = base + (uint32_t)(e - Q.set_ks0);
         k.row0 = k.row1 = (uint32_t)r1;
         k.l = 0;
         for (size_t r = r0; r < r1; r++) {
            const RowRec &row = Q.rows[r];
            const uint32_t nlabcd = (row.nlnp & 15) * ((row.nlnp >> 4) & 15) * ((row.nlnp >> 8) & 15) * ((row.nlnp >> 12) & 15);
            if (q >= row.q_off && q < row.q_off + nlabcd) {
               size_t r2 = r;
               while (r2 < r1 && Q.rows[r2].q_off == row.q_off) r2++;
               k.row0 = (uint32_t)r;
               k.row1 = (uint32_t)r2;
               k.l = q - row.q_off;
               break;
            }
         }
         if (q >= base + Q.set_nq) k.row0 = k.row1 = (uint32_t)r1;   // no integrals declared for it
      }
      Q.n_q += Q.set_nq;
      Q.touched = false;

      // memory of the pending batch of this class
      const ClassPlan cp = make_plan(c / 64, (c / 16) % 4, (c / 4) % 4, c % 4);
      const size_t idx_bytes = Q.rows.size() * sizeof(RowRec) + Q.prm_ip.size() * 8 + Q.ks.size() * sizeof(KsRec);
      const size_t eco_bytes = (size_t)Q.e_slots * cp.ncp * sizeof(double);
      const size_t min_ac0 = (size_t)64 * cp.ncp * sizeof(double);
      if (idx_bytes + eco_bytes + min_ac0 > budget_bytes_ / 2) {
         const size_t set_idx = (Q.rows.size() - r0) * sizeof(RowRec) + (Q.prm_ip.size() - Q.set_prm0) * 8 +
                                (Q.ks.size() - Q.set_ks0) * sizeof(KsRec);
         const size_t set_eco = (size_t)(Q.e_slots - Q.set_eslots0) * cp.ncp * sizeof(double);
         if (r0 > 0 && set_idx + set_eco + min_ac0 <= budget_bytes_ / 2) {
            dispatch_class(c, r0, Q.set_prm0, Q.set_ks0);   // earlier sets only
         } else {
            dispatch_class(c, Q.rows.size(), Q.prm_ip.size(), Q.ks.size());
         }
      }
   }
   set_classes_.clear();
}

void Engine::dispatch_all()
{
   if (!set_classes_.empty()) fatal("dispatch with an open set: call add_set first", (int)set_classes_.size());
   for (int c = 0; c < NCLASS; c++)
      if (!q_[c].ks.empty() || !q_[c].rows.empty()) dispatch_class(c, q_[c].rows.size(), q_[c].prm_ip.size(), q_[c].ks.size());
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

void Engine::ensure_arena()
{
   if (d_arena_ && arena_bytes_ == budget_bytes_) return;
   if (d_arena_) {
      synchronize();
      HIPCHECK(hipFree(d_arena_));
      HIPCHECK(hipHostFree(h_stage_));
   }
   arena_bytes_ = budget_bytes_;
   HIPCHECK(hipMalloc((void **)&d_arena_, arena_bytes_));
   stage_bytes_ = arena_bytes_ / 2;
   HIPCHECK(hipHostMalloc((void **)&h_stage_, stage_bytes_, hipHostMallocDefault));
   if (!d_stats_) {
      HIPCHECK(hipMalloc((void **)&d_stats_, 4 * sizeof(unsigned long long)));
      HIPCHECK(hipMemset(d_stats_, 0, 4 * sizeof(unsigned long long)));
   }
   for (auto &e : ev_)
      if (!e) HIPCHECK(hipEventCreate((hipEvent_t *)&e));
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

void Engine::dispatch_class(int c, size_t nrows, size_t nprm, size_t nks)
{
   ClassQueue &Q = q_[c];
   if (nrows == 0 && nks == 0) return;
   if (!d_P_[0]) fatal("integrals dispatched before set_P", c);
   ensure_arena();
   ensure_tables();
   upload_data();
   hipStream_t st = (hipStream_t)stream_;
   const ClassPlan cp = make_plan(c / 64, (c / 16) % 4, (c / 4) % 4, c % 4);

   // quartets and e-slots of the prefix
   uint32_t nq = 0, eslots = 0;
   if (nrows < Q.rows.size()) {
      nq = Q.rows[nrows].q_off;
      eslots = Q.rows[nrows].e_off;
   } else {
      nq = Q.n_q;
      eslots = Q.e_slots;
   }
   // index data -> pinned staging -> arena
   const size_t b_rows = align256(nrows * sizeof(RowRec));
   const size_t b_prow = align256(nprm * 4), b_pip = align256(nprm * 4);
   const size_t b_ks = align256(nks * sizeof(KsRec));
   const size_t idx_bytes = b_rows + b_prow + b_pip + b_ks;
   const size_t eco_bytes = align256((size_t)eslots * cp.ncp * sizeof(double));
   if (idx_bytes > stage_bytes_ || idx_bytes + eco_bytes + (size_t)64 * cp.ncp * 8 > arena_bytes_)
      fatal("one set quartet needs more device memory than set_max_mem allows", (int)((idx_bytes + eco_bytes) >> 20));
   HIPCHECK(hipStreamSynchronize(st));   // staging and arena are free again
   char *hs = h_stage_;
   std::memcpy(hs, Q.rows.data(), nrows * sizeof(RowRec));
   std::memcpy(hs + b_rows, Q.prm_row.data(), nprm * 4);
   std::memcpy(hs + b_rows + b_prow, Q.prm_ip.data(), nprm * 4);
   std::memcpy(hs + b_rows + b_prow + b_pip, Q.ks.data(), nks * sizeof(KsRec));
   HIPCHECK(hipMemcpyAsync(d_arena_, h_stage_, idx_bytes, hipMemcpyHostToDevice, st));

   BatchDev b;
   b.rows = (const RowRec *)d_arena_;
   b.prm_row = (const uint32_t *)(d_arena_ + b_rows);
   b.prm_ip = (const uint32_t *)(d_arena_ + b_rows + b_prow);
   b.ks = (const KsRec *)(d_arena_ + b_rows + b_prow + b_pip);
   b.data = d_data_;
   b.eco = (double *)(d_arena_ + idx_bytes);
   b.ac0 = (double *)(d_arena_ + idx_bytes + eco_bytes);
   b.n_rows = (uint32_t)nrows;
   b.n_prm = (uint32_t)nprm;
   b.n_q = (uint32_t)nks;
   b.stats = stats_on_ ? d_stats_ : nullptr;
   const size_t ac0_cap = (arena_bytes_ - idx_bytes - eco_bytes) / (sizeof(double) * cp.ncp);

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

   const size_t need = scratch_doubles(cp);
   if (need > scratch_cap_) {
      if (d_scratch_) HIPCHECK(hipFree(d_scratch_));
      HIPCHECK(hipMalloc((void **)&d_scratch_, need * sizeof(double)));
      scratch_cap_ = need;
   }
   Scratch sc{d_scratch_, scratch_cap_};

   if (eco_bytes) HIPCHECK(hipMemsetAsync(b.eco, 0, eco_bytes, st));
   const bool tm = timing_on_;
   // primitives in chunks that fit the arena; the contraction accumulates
   for (size_t p0 = 0; p0 < nprm; p0 += ac0_cap) {
      const size_t p1 = std::min(nprm, p0 + ac0_cap);
      BatchDev bc = b;
      bc.prm_row = b.prm_row + p0;
      bc.prm_ip = b.prm_ip + p0;
      bc.n_prm = (uint32_t)(p1 - p0);
      if (tm) HIPCHECK(hipEventRecord((hipEvent_t)ev_[0], st));
      int rc = launch_fm_vrr(cp, t, bc, sc, st);
      if (rc) fatal("fm_vrr launch", rc);
      if (tm) HIPCHECK(hipEventRecord((hipEvent_t)ev_[1], st));
      // rows whose primitives intersect [p0, p1)
      const uint32_t first_row = Q.prm_row[p0];
      const uint32_t last_row = Q.prm_row[p1 - 1];
      rc = launch_eco_range(cp, b, first_row, last_row + 1, (uint32_t)p0, (uint32_t)p1, st);
      if (rc) fatal("eco launch", rc);
      if (tm) {
         HIPCHECK(hipEventRecord((hipEvent_t)ev_[2], st));
         HIPCHECK(hipEventSynchronize((hipEvent_t)ev_[2]));
         float ms0 = 0, ms1 = 0;
         HIPCHECK(hipEventElapsedTime(&ms0, (hipEvent_t)ev_[0], (hipEvent_t)ev_[1]));
         HIPCHECK(hipEventElapsedTime(&ms1, (hipEvent_t)ev_[1], (hipEvent_t)ev_[2]));
         timing_.ms[0] += ms0;
         timing_.ms[1] += ms1;
         timing_.launches[0]++;
         timing_.launches[1]++;
      }
   }
   KTargets kt;
   kt.nspin = nspin_;
   kt.hf_fac = hf_fac_;
   for (int s = 0; s < 2; s++) {
      kt.P[s] = d_P_[s];
      kt.K[s] = d_K_[s];
   }
   if (tm) HIPCHECK(hipEventRecord((hipEvent_t)ev_[3], st));
   int rc = launch_hrr_sph_ks(cp, t, b, kt, sc, st);
   if (rc) fatal("hrr_sph_ks launch", rc);
   if (tm) {
      HIPCHECK(hipEventRecord((hipEvent_t)ev_[4], st));
      HIPCHECK(hipEventSynchronize((hipEvent_t)ev_[4]));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, (hipEvent_t)ev_[3], (hipEvent_t)ev_[4]));
      timing_.ms[2] += ms;
      timing_.launches[2]++;
   }
   {  // algorithmic FLOPs of the contraction and of the quartet kernel
      double eco = 0, qk = 0;
      for (size_t r = 0; r < nrows; r++) {
         const RowRec &row = Q.rows[r];
         const double nlabcd = (double)((row.nlnp & 15) * ((row.nlnp >> 4) & 15) * ((row.nlnp >> 8) & 15) * ((row.nlnp >> 12) & 15));
         eco += 2.0 * nlabcd * cp.ncp * row.n_prm;
         qk += nlabcd * cp.hrr_flops;
      }
      qk += (double)nks * (cp.sph_flops + 9.0 * cp.Ns);
      timing_.flops[1] += eco;
      timing_.flops[2] += qk;
   }
   n_slots_ += (uint64_t)nprm * (uint64_t)t.Ng;
   n_prm_ += nprm;
   n_q_ += nks;
   n_rows_ += nrows;

   // drop the dispatched prefix, rebase the rest
   Q.rows.erase(Q.rows.begin(), Q.rows.begin() + (long)nrows);
   Q.prm_row.erase(Q.prm_row.begin(), Q.prm_row.begin() + (long)nprm);
   Q.prm_ip.erase(Q.prm_ip.begin(), Q.prm_ip.begin() + (long)nprm);
   Q.ks.erase(Q.ks.begin(), Q.ks.begin() + (long)nks);
   for (RowRec &r : Q.rows) {
      r.prm_off -= (uint32_t)nprm;
      r.q_off -= nq;
      r.e_off -= eslots;
   }
   for (uint32_t &r : Q.prm_row) r -= (uint32_t)nrows;
   for (KsRec &k : Q.ks) {
      k.row0 -= (uint32_t)nrows;
      k.row1 -= (uint32_t)nrows;
   }
   Q.n_q -= nq;
   Q.e_slots -= eslots;
   if (Q.touched) {
      Q.set_row0 -= nrows;
      Q.set_prm0 -= nprm;
      Q.set_ks0 -= nks;
      Q.set_eslots0 -= eslots;
   }
}

void Engine::enable_stats(bool on) { stats_on_ = on; }
void Engine::enable_timing(bool on) { timing_on_ = on; }

void Engine::reset_counters()
{
   synchronize();
   n_slots_ = n_prm_ = n_q_ = n_rows_ = 0;
   timing_ = Timing();
   if (d_stats_) HIPCHECK(hipMemset(d_stats_, 0, 4 * sizeof(unsigned long long)));
}

void Engine::stats(uint64_t out[8])
{
   synchronize();
   unsigned long long s[4] = {0, 0, 0, 0};
   if (d_stats_) HIPCHECK(hipMemcpy(s, d_stats_, sizeof s, hipMemcpyDeviceToHost));
   out[0] = n_slots_;
   out[1] = s[0];
   out[2] = n_prm_;
   out[3] = n_q_;
   out[4] = s[1];
   out[5] = n_rows_;
   out[6] = s[2];
   out[7] = 0;
}

}  // namespace gint
