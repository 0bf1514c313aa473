// Host engine behind the libgint_* C-ABI (one object per handle, i.e. per
// OpenMP thread of the caller).  Mirrors the state and call protocol of the
// reference class libGint (reference src/libGint.h:53-229,
// src/libGint.cpp), with its own batching and device layout.
#pragma once
#include <stdint.h>

#include <unordered_map>
#include <vector>

#include "device.hpp"

namespace gint {

// Append-only store of basis data (positions, exponents, coefficients) with
// exact deduplication: equal arrays share one copy (hash then compare; the
// reference's src/UniqueArray.cpp:60-92 trusts the hash alone).
class DataStore {
 public:
   uint32_t add(const double *v, int n);
   const std::vector<double> &data() const { return buf_; }
   uint64_t version() const { return version_; }

 private:
   std::vector<double> buf_;
   std::unordered_multimap<uint64_t, std::pair<uint32_t, int>> index_;
   uint64_t version_ = 0;
};

struct SetInfo {
   uint32_t R = 0, Z = 0;
   int np = 0;
   std::vector<int> l, nl;
   std::vector<uint32_t> K;
};

// pending work of one angular-momentum class
struct ClassQueue {
   std::vector<RowRec> rows;
   std::vector<uint32_t> prm_row, prm_ip;
   std::vector<KsRec> ks;
   uint32_t n_q = 0;       // contracted quartets committed
   uint32_t e_slots = 0;   // sum of nlabcd over rows
   // current set
   size_t set_row0 = 0, set_prm0 = 0, set_ks0 = 0;
   uint32_t set_nq = 0, set_eslots0 = 0;
   bool touched = false;
};

struct Timing {
   double ms[3] = {0, 0, 0};   // fm_vrr, eco, hrr_sph_ks
   uint64_t launches[3] = {0, 0, 0};
   double flops[3] = {0, 0, 0};  // algorithmic FLOPs of eco and hrr_sph_ks (host-counted); [0] from stats
};

class Engine {
 public:
   Engine();
   ~Engine();
   void init();
   void set_potential_truncated(double R_cut, const double *C0, int ld_C0, int C0_size);
   void set_hf_fac(double f) { hf_fac_ = f; }
   void set_max_mem(int max_mem_MiB);
   void set_P(const double *P, int n);
   void set_P(const double *Pa, const double *Pb, int n);
   void set_K(const double *K, int n);
   void set_K(const double *Ka, const double *Kb, int n);
   void get_K(double *K);
   void get_K(double *Ka, double *Kb);
   void set_cell(const double *h, const double *hinv);
   void set_neighs(const double *g, int n);
   void set_atom(int i, const double *R, const double *Z, int np);
   void set_atom_L(int i, int l, int nl, const double *K_normalised);
   void add_prm(int ipa, int ipb, int ipc, int ipd);
   void add_shell(int i, int j, int k, int l, int n1, int n2);
   void add_qrt(int la, int lb, int lc, int ld, int nla, int nlb, int nlc, int nld);
   void add_qrtt(double symm_fac, int la, int lb, int lc, int ld, int inla, int inlb, int inlc, int inld,
                 int ld_ac, int ld_ad, int ld_bc, int ld_bd, unsigned off_ac, unsigned off_ad, unsigned off_bc,
                 unsigned off_bd, bool Tac, bool Tad, bool Tbc, bool Tbd);
   void add_set();
   void dispatch_all();
   void synchronize();

   // extensions (not in the reference ABI)
   void enable_stats(bool on);
   void stats(uint64_t out[8]);
   void enable_timing(bool on);
   Timing timing() const { return timing_; }
   void reset_counters();
   int np_of(int i) const { return i < (int)sets_.size() ? sets_[i].np : 0; }
   const double *exponents_of(int i) const { return &store_.data()[sets_[i].Z]; }

 private:
   void dispatch_class(int c, size_t nrows, size_t nprm, size_t nks);
   void ensure_arena();
   void ensure_tables();
   void upload_data();
   void free_device_matrices();
   int my_thr_ = 0, nomp_ = 1;
   void *stream_ = nullptr;

   // geometry, potential
   double cell_[18];
   std::vector<double> neighs_;
   bool truncated_ = false;
   double R_cut_ = 0.0;
   std::vector<double> C0_;
   int ld_C0_ = 0;
   double hf_fac_ = 1.0;
   bool tables_dirty_ = true;

   // basis
   DataStore store_;
   std::vector<SetInfo> sets_;
   uint64_t data_version_dev_ = ~0ull;

   // work
   std::vector<uint32_t> prm_tmp_;
   ClassQueue q_[NCLASS];
   std::vector<int> set_classes_;
   size_t budget_bytes_ = 0;

   // device
   double *d_data_ = nullptr;
   size_t d_data_cap_ = 0;
   double *d_C0_ = nullptr;
   int *d_grid_ = nullptr;
   double *d_bw_ = nullptr, *d_ftab_ = nullptr, *d_cell_ = nullptr, *d_neighs_ = nullptr;
   size_t d_neighs_cap_ = 0;
   char *d_arena_ = nullptr;
   size_t arena_bytes_ = 0;
   char *h_stage_ = nullptr;
   size_t stage_bytes_ = 0;
   double *d_scratch_ = nullptr;
   size_t scratch_cap_ = 0;
   unsigned long long *d_stats_ = nullptr;
   bool stats_on_ = false;
   uint64_t n_slots_ = 0, n_prm_ = 0, n_q_ = 0, n_rows_ = 0;

   // P and K: shared by the team (reference src/libGint.cpp:441-447,558-562)
   int nspin_ = 0;
   size_t FP_size_ = 0;
   double *d_P_[2] = {nullptr, nullptr};
   double *d_K_[2] = {nullptr, nullptr};

   bool timing_on_ = false;
   Timing timing_;
   void *ev_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

void fatal(const char *what, int code);

}  // namespace gint
