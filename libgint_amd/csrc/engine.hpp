// Host engine behind the libgint_* C-ABI (one object per handle, i.e. per
// OpenMP thread of the caller).  Mirrors the state and call protocol of the
// reference class libGint (reference src/libGint.h:53-229,
// src/libGint.cpp), with its own batching and device layout:
//
//  * work is grouped by quartet type (shell structure of the four sets);
//    one row per add_shell call keeps its primitive list once for every
//    class of the set quartet;
//  * the add_qrtt calls of a (set quartet, class) are compressed to offsets
//    affine in the contraction indices (or kept as an explicit list);
//  * pending work is uploaded and run when it exceeds the memory budget or
//    at dispatch / get_K, or, in plan mode, kept resident on the device and
//    run again for every new density.
#pragma once
#include <stdint.h>

#include <array>
#include <cstring>
#include <map>
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
   int sig = -1;   // shell-structure id, -1 until needed
};

// host-side pending work of one quartet type
struct TypeQueue {
   std::vector<RowRec> rows;
   std::vector<uint32_t> prm_ip, prm_row;
   std::vector<GroupRec> groups;
   std::vector<KsGroup> ksg;
   std::vector<KsSet> kss;
   std::vector<KsRec> ksx;
   bool empty() const { return groups.empty(); }
   void release();
};

struct Chunk {
   uint32_t g0, g1, r0, r1, p0, p1;
   double tq_flops, k3_flops, eco_flops;
};

// one type's work on the device
struct TypeDev {
   int type = -1;
   RowRec *rows = nullptr;
   GroupRec *groups = nullptr;
   uint32_t *prm_ip = nullptr, *prm_row = nullptr;
   KsGroup *ksg = nullptr;
   KsSet *kss = nullptr;
   KsRec *ksx = nullptr;
   size_t cap[7] = {0, 0, 0, 0, 0, 0, 0};   // allocated elements of each array
   size_t n_rows = 0, n_groups = 0, n_prm = 0, n_ksx = 0, n_ksg = 0, n_kss = 0;
   size_t max_group_bytes = 0;            // largest set quartet's primitive blocks
   size_t max_chunk_bytes = 0;            // largest chunk's blocks and contraction buffer rows
   std::vector<Chunk> chunks;
   std::vector<Chunk> chunks_na;          // the same work in chunks that fit the no-allocation scratch
   bool na_fit = true;                    // every set quartet fits it
   std::vector<char> class_used;
   double *gmag = nullptr;      // mixed precision: calibrated max |block element| per set quartet
   uint32_t *gflag = nullptr;   // mixed precision: route per set quartet (1 single precision)
   void release();
};

struct TypeInfo {
   std::array<int, 4> sig;
   TypePlan plan;
   TypePlan *d_plan = nullptr;
   std::vector<uint32_t> e1, e2;          // Hermite -> cartesian item descriptors
   uint32_t *d_e1 = nullptr, *d_e2 = nullptr;
   std::vector<ClassPlan> cplans;
   std::vector<CqLayout> lays;            // quartet kernel: per-class LDS plans
   std::vector<int> srcmaps;              // class block element -> type block element
   int cq_lds = 0, cq_gws = 0;
   bool cq_wide = false;   // a class takes the Kronecker path: k_gqs with four wavefronts
   ClassPlan *d_cps = nullptr;
   CqLayout *d_lays = nullptr;
   int *d_srcmaps = nullptr;
   // matrix-core contraction (k_eco): work items of a row, [class map
   // offsets, class buffer offsets], doubles per row of the buffer
   std::vector<EcoItem> eco_items;
   std::vector<int> eco_info;
   size_t eco_row = 0;
   EcoItem *d_eco_items = nullptr;
   int *d_eco_info = nullptr;
};

// device buffers of one chunk's kernels
struct RunBufs {
   double *ac0;
   uint32_t *fix, *fixcnt;
   double *fix_ws;
   uint32_t fix_cap;
   unsigned long long *stats;
   Scratch gws;
   bool count;   // host counters and timing events
};

struct Timing {
   double ms[5] = {0, 0, 0, 0, 0};        // fm_r, tq, gqs, fixup, eco (matrix-core contraction)
   uint64_t launches[5] = {0, 0, 0, 0, 0};
   double flops[5] = {0, 0, 0, 0, 0};     // host-counted algorithmic FLOPs of tq [1], gqs [2] and eco [4]
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
   // extension: every add_qrt / add_qrtt call of the current set quartet at
   // once, for K blocks laid out per pair of sets (KsSet)
   void add_set_blocks(double symm_fac, const int ld[4], const unsigned off[4], const bool T[4]);
   // extension: integral output (the reference's TRA layout) in place of K
   void set_output(double *out, uint64_t n);
   void add_tra(int la, int lb, int lc, int ld, uint64_t dest);
   void get_output();
   void dispatch_all();
   void synchronize();

   // extensions (not in the reference ABI)
   void enable_stats(bool on) { stats_on_ = on; }
   void stats(uint64_t out[8]);
   void enable_timing(bool on) { timing_on_ = on; }
   Timing timing();
   void timing_detail(int cat, double out[256]);
   void reset_counters();
   void set_debug(int flags) { debug_ = flags; }
   void force_kron(std::vector<TypeDev> &devs, bool own_first);
   void set_graph(bool on) { graph_on_ = on; }
   void set_autotune(bool on) { autotune_on_ = on; }
   void set_mixed_precision(double tau) { mixed_tau_ = tau > 0.0 ? tau : 0.0; }
   void mixed_info(double out[4]);
   void tune_info(double out[5]) const
   {
      out[0] = plan_tuned_ ? 1.0 : 0.0;
      for (int i = 0; i < 3; i++) out[1 + i] = (double)tune_counts_[i];
      out[3] += (double)tune_counts_[3];
      out[2] += (double)tune_counts_[4];
      out[4] = tune_seconds_;
   }
   // per class code: ms of the chosen schedule on the timing samples, and the schedule
   void tune_detail(double ms[256], int pick[256]) const
   {
      for (int i = 0; i < 256; i++) {
         ms[i] = tune_ms_[i];
         pick[i] = tune_pick_[i];
      }
   }
   void graph_info(uint64_t out[4]) const
   {
      out[0] = graph_on_ ? 1 : 0;
      out[1] = graph_captures_;
      out[2] = graph_replays_;
      out[3] = graph_exec_ ? 1 : 0;
   }
   void plan_begin();
   void plan_end();
   void plan_execute();
   // extension: the resident plan into caller-owned device P / K, with every
   // scratch buffer in the library's static no-allocation array; 0 or < 0
   int plan_execute_noalloc(const double *Pa, const double *Pb, double *Ka, double *Kb, uint64_t n);
   void plan_clear();
   void plan_info(uint64_t out[6]);
   // serving (server.cpp)
   void use_private_stream();
   void *stream() const { return stream_; }
   uint64_t device_bytes() const;       // device memory held now (plan, work buffers, tables, basis)
   uint64_t plan_host_bytes() const;    // the evicted plan's host copy
   uint64_t work_estimate() const;      // work buffers an execution of the resident plan allocates
   uint64_t plan_max_index() const { return plan_max_index_; }
   bool plan_evicted() const { return evicted_; }
   void plan_evict();                   // plan records and work buffers off the device, host copy kept
   void plan_restore();
   void plan_execute_into(const double *P0, const double *P1, double *K0, double *K1, int nspin, uint64_t n);
   void queue_info(uint64_t out[8]);
   int queue_types(uint64_t *out, int max_types);
   int np_of(int i) const { return i < (int)sets_.size() ? sets_[i].np : 0; }
   const double *exponents_of(int i) const { return &store_.data()[sets_[i].Z]; }

 private:
   int type_of(int i, int j, int k, int l);
   int set_sig(int i);
   KsGroup compress(std::vector<KsRec> &calls, const TypePlan &tp, int k, std::vector<KsRec> &ksx);
   bool set_level(const KsGroup *g, const TypePlan &tp, KsSet &s) const;
   void upload(std::vector<TypeQueue> &qs, std::vector<TypeDev> &devs, bool release_host);
   void execute(std::vector<TypeDev> &devs, uint64_t max_index);
   bool prepare_exec(std::vector<TypeDev> &devs, uint64_t max_index, RunBufs &w, KTargets &kt, uint64_t n);
   void launch_all(std::vector<TypeDev> &devs, const RunBufs &w, const KTargets &kt);
   void replay_graph(const RunBufs &w, const KTargets &kt);
   void autotune(const RunBufs &w, const KTargets &kt);
   void apply_schedule(TypeInfo &ti, int k, int v);
   void upload_layouts(TypeInfo &ti);
   void drop_graph();
   static uint64_t graph_bits(double x)
   {
      uint64_t u;
      std::memcpy(&u, &x, 8);
      return u;
   }
   void run_chunk(const TypeInfo &ti, const TypeDev &td, const Chunk &c, const RunBufs &w, const KTargets &kt);
   DevTables dev_tables() const;
   void ensure_work(size_t ncp_max, size_t max_prm_chunk, size_t max_group_bytes);
   void ensure_tables();
   void upload_data();
   void harvest_events();
   void ev_begin(int cat, void *stream);
   void ev_end(int cat, void *stream);
   void ensure_side();
   int my_thr_ = 0, nomp_ = 1;
   void *stream_ = nullptr;
   // the consumers of a chunk's blocks only add into K: the small-class
   // kernels run on side streams beside gqs on stream_, and
   // stream_ waits for them before the blocks are overwritten
   static constexpr int NSIDE = 3;
   void *side_[NSIDE] = {};
   void *ev_join_[NSIDE + 1] = {};   // [0] blocks ready, [1 + i] side stream i done

   // geometry, potential
   double cell_[18];
   std::vector<double> neighs_;
   bool truncated_ = false;
   double R_cut_ = 0.0;
   std::vector<double> C0_;
   int ld_C0_ = 0;
   int nder_ = -1;   // highest derivative order of the C0 table
   double hf_fac_ = 1.0;
   bool tables_dirty_ = true;

   // basis
   DataStore store_;
   std::vector<SetInfo> sets_;
   std::map<std::pair<std::vector<int>, std::vector<int>>, int> sig_ids_;
   std::map<std::array<int, 4>, int> type_ids_;
   std::vector<TypeInfo> types_;
   uint64_t data_version_dev_ = ~0ull;
   std::array<int, 4> last_sets_ = {-1, -1, -1, -1};
   int last_type_ = -1;

   // ingestion
   std::vector<uint32_t> prm_tmp_;
   std::vector<TypeQueue> q_;
   int cur_type_ = -1;
   std::array<int, 4> cur_sets_ = {-1, -1, -1, -1};
   std::vector<KsRec> pend_[NCLASS];
   std::vector<int> touched_;
   KsSet blocks_{};            // add_set_blocks of the open set quartet
   bool have_blocks_ = false;
   std::vector<std::pair<int, uint64_t>> pend_out_;   // (class code, destination) of the open set quartet
   double *h_out_ = nullptr;   // caller's output array (set_output)
   double *d_out_ = nullptr;
   uint64_t out_n_ = 0;
   size_t pend_bytes_ = 0;
   uint64_t max_index_ = 0;   // highest P / K element the pending calls reach
   uint64_t plan_max_index_ = 0;
   size_t budget_bytes_ = 0;
   bool recording_ = false;
   int debug_ = 0;

   // device
   std::vector<TypeDev> stream_dev_, plan_dev_;
   double *d_data_ = nullptr;
   size_t d_data_cap_ = 0;
   double *d_C0_ = nullptr;
   int *d_grid_ = nullptr;
   double *d_bw_ = nullptr, *d_ftab_ = nullptr, *d_cell_ = nullptr, *d_neighs_ = nullptr;
   size_t d_neighs_cap_ = 0;
   double *d_work_ = nullptr;   // ac0 + eco of one chunk
   size_t work_bytes_ = 0;
   uint32_t *d_fix_ = nullptr;
   double *d_fix_ws_ = nullptr;
   size_t fix_cap_ = 0, fix_ws_cap_ = 0;
   double *d_scratch_ = nullptr;
   size_t scratch_cap_ = 0;
   void *noalloc_scratch_ = nullptr;   // this device's static scratch array (init)
   unsigned long long *d_stats_ = nullptr;   // [0..2] kernel statistics, [3] unused
   uint32_t *d_fixcnt_ = nullptr;             // [0] entries of the chunk, [1] slots beyond the list (sticky)
   uint64_t fix_overflow_ = 0;                // such slots seen so far (kept by every class)
   bool stats_on_ = false;
   uint64_t n_slots_ = 0, n_prmc_ = 0, n_q_ = 0, n_rows_ = 0;

   // P and K: shared by the team (reference src/libGint.cpp:441-447,558-562)
   int nspin_ = 0;
   size_t FP_size_ = 0;
   double *d_P_[2] = {nullptr, nullptr};
   double *d_K_[2] = {nullptr, nullptr};

   // autotuner: the schedule of each class, chosen at the plan's first execution
   static constexpr uint32_t TUNE_GROUPS = 16384;  // at most this many set quartets in a type's timing sample
   static constexpr double TUNE_PRM = 262144;       // primitive quartets aimed at in the sample
   static constexpr double TUNE_VALU_FLOPS = 2e8;   // largest sample contraction timed on the VALU
   bool autotune_on_ = true;
   bool plan_tuned_ = false;
   // classes given k_tqw / k_gqs / k_eco + k_gqs / k_eco + Kronecker k_gqs / Kronecker k_gqs with its own contraction
   uint64_t tune_counts_[5] = {0, 0, 0, 0, 0};
   double tune_seconds_ = 0.0;
   double tune_ms_[256] = {};
   int tune_pick_[256] = {};

   // mixed precision (plans): set quartets whose calibrated bound times
   // max |P| is below tau take the single-precision contraction
   double mixed_tau_ = 0.0;
   bool mixed_calibrated_ = false;
   int mixed_mode_ = 0;                    // of the execution being launched: 0 off, 1 calibrate, 2 route
   unsigned long long *d_mix_ = nullptr;   // [0] max |P| bits, [1] set quartets routed to single precision
   void ensure_mixed();

   bool own_stream_ = false;
   bool evicted_ = false;
   std::vector<TypeQueue> plan_host_;

   // the resident plan as a HIP graph (plan_execute)
   bool graph_on_ = true;
   void *graph_exec_ = nullptr;
   uint64_t graph_key_[32] = {};
   uint64_t graph_counts_[2] = {0, 0};
   double graph_flops_[3] = {0, 0, 0};
   uint64_t graph_captures_ = 0, graph_replays_ = 0;

   bool timing_on_ = false;
   Timing timing_;
   struct EvRec {
      int cat;
      void *e0, *e1;
      int key;
   };
   int ev_key_ = 0;
   std::array<std::array<double, 256>, 5> detail_{};   // ms per category and key (type L / class code)
   std::vector<void *> ev_pool_;
   std::vector<EvRec> ev_live_;
   size_t ev_used_ = 0;
   void *ev_open_ = nullptr;
};

void fatal(const char *what, long long code);

}  // namespace gint
