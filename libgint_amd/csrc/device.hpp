// Interface between the host engine (engine.cpp, built with g++ -fopenmp)
// and the gfx950 kernels (kernels.hip, built with hipcc): record layouts in
// HBM, the per-type and per-class plans, and the launchers.  Plain types
// only.
//
// Work is organised by quartet *type*: the shell structure (l and number of
// contractions of every shell) of the four sets of a set quartet.  All
// shells of a set share its exponents, so the primitive work of a
// (primitive quartet, lattice cell) slot is done once for every
// angular-momentum class of the set quartet.
#pragma once
#include <stdint.h>
#include <stddef.h>

#include <vector>

#if defined(__HIPCC__)
#define GINT_HDI __host__ __device__ inline
#else
#define GINT_HDI inline
#endif

namespace gint {

constexpr int MAXL = 3;          // highest angular momentum of a shell (f)
constexpr int MAXLL = 4 * MAXL;  // highest total angular momentum of a class
constexpr int MAXNL = 4;         // contractions per shell
constexpr int MAXNP = 16;        // primitives per set
constexpr int MAXNG = 256;       // lattice vectors
constexpr int NCLASS = 256;      // (la,lb,lc,ld), each 0..3
constexpr int MAXSH = 4;         // shells per set (distinct l)
constexpr int MAXTC = MAXSH * MAXSH * MAXSH * MAXSH;  // classes per type
constexpr int MAXHERM = 455;     // Hermite triples t+u+v <= 12

inline constexpr int nco(int l) { return (l + 1) * (l + 2) / 2; }
inline constexpr int nsph(int l) { return 2 * l + 1; }
inline constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
inline constexpr int class_code(int la, int lb, int lc, int ld) { return ((la * 4 + lb) * 4 + lc) * 4 + ld; }

// One row: a set quartet at bra/ket images (n1,n2) with its primitive
// quartets (one libgint_add_shell call).  Its contracted blocks sit at
// (row - first row of the chunk) * TypePlan::ebytes in the contraction buffer.
struct RowRec {
   uint32_t prm_off;   // first primitive quartet of the row
   uint32_t n_prm;
   uint32_t group;     // its set quartet
   uint32_t n12;       // n1 | n2 << 16
};
static_assert(sizeof(RowRec) == 16, "RowRec layout");

// One set quartet (the work between two libgint_add_set calls).
struct GroupRec {
   uint32_t row0, row1;       // its rows
   uint32_t R[4], Z[4];       // centre and exponent offsets in the basis data
   uint32_t Kof[4][MAXSH];    // normalised coefficients of shell s of centre t (nl x np)
   uint32_t np4;              // np-1 of the four sets, 4-bit fields
   uint32_t set_ij, set_kl;   // set indices, 16-bit fields (survivor checksum only)
   uint32_t ks;               // KS_SETLEVEL: its KsSet; else its first KsGroup (nclass of them)
   uint32_t ksmode;           // KS_SETLEVEL or 0
   uint32_t pad;
};
static_assert(sizeof(GroupRec) == 128, "GroupRec layout");

// Where the contracted quartets of one (set quartet, class) go in K: the
// add_qrtt calls of the class compressed to offsets affine in the
// contraction indices (inla, inlb, inlc, inld), or a reference to an
// explicit list when the caller's offsets are not affine.
struct KsGroup {
   uint32_t base[4];   // ac ad bc bd block offsets of the first call
   int32_t step[8];    // ac: by inla, inlc; ad: by inla, inld; bc: by inlb, inlc; bd: by inlb, inld
   uint32_t ld01;      // ld_ac | ld_ad << 16
   uint32_t ld23;      // ld_bc | ld_bd << 16
   uint32_t flags;     // bits 0-3 transposes, 4-6 symmetry code, bit 8 explicit, bits 16-24 calls
   uint32_t expl;      // first explicit record when bit 8 is set
};
static_assert(sizeof(KsGroup) == 64, "KsGroup layout");

constexpr uint32_t KS_EXPLICIT = 1u << 8;
// integral output instead of a K update: the class's spherical integrals go
// to the output array at base[0] | base[1] << 32, ordered
// [ila, isa, ilb, isb, ilc, isc, ild, isd] (reference src/compute_TRA.cpp:36-47)
constexpr uint32_t KS_OUTPUT = 1u << 9;
GINT_HDI uint32_t ks_count(uint32_t flags) { return flags >> 16; }

// Every add_qrtt call of one set quartet at once, when the caller lays out
// its P / K blocks per pair of sets (the layout of CP2K's block matrices):
// function f of set x and function g of set y sit at
// base + (T ? f ld + g : g ld + f) of the (x,y) pair block, with the
// functions of a set ordered shell by shell, contraction by contraction,
// 2l+1 each.  Every class of the set quartet has all its contraction
// combinations, and they share leading dimensions, transposes and the
// symmetry code.  The KsGroup of each class follows from it (ks_from_set).
struct KsSet {
   uint32_t base[4];   // ac ad bc bd pair blocks: element of (function 0, function 0)
   uint32_t ld01;      // ld_ac | ld_ad << 16
   uint32_t ld23;      // ld_bc | ld_bd << 16
   uint32_t flags;     // bits 0-3 transposes, 4-6 symmetry code
   uint32_t pad;
};
static_assert(sizeof(KsSet) == 32, "KsSet layout");

constexpr uint32_t KS_SETLEVEL = 1u;   // GroupRec::ksmode

// one add_qrtt call, kept when a (set quartet, class) is not affine
struct KsRec {
   uint32_t off[4];
   uint32_t ld01, ld23, flags, pad;
};
static_assert(sizeof(KsRec) == 32, "KsRec layout");

// Plan of one angular-momentum class: HRR layers, cart->sph, sizes.
struct ClassPlan {
   int la, lb, lc, ld;
   int L, LA, LC;
   int ncp;                  // its (e0|f0) block set, e in [la,LA], f in [lc,LC], e outer
   int cpoff[7][7];
   int Nc, Ns;
   int koff[4][7][7];        // ket HRR layer dd: block (e,c), e in [la,LA], c in [lc,LC-dd]
   int ksize[4];
   int boff[4][7];           // bra HRR layer bb: block a in [la, LA-bb]
   int bsize[4];
   int ycap;                 // capacity of one layer buffer (>= ncp, Nc, every layer)
   int hrr_ws;               // HRR workspace doubles per workgroup: 3 ycap + Nc
   // algorithmic FLOPs of the reference chain for this class
   int fm_flops;             // 52 + 211 (L+1) per surviving slot
   int vrr_flops;            // reference plan VRR (weights 3/8/3/8/11/6 per element)
   int hrr_flops;            // 2 per HRR element
   int sph_flops;            // 2 m n k over the four stages
};

// Plan of one quartet type.  Blocks (e0|f0) with e in [LA0,LA], f in
// [LC0,LC] (e outer, layout [ie][jf] inside a block) are produced per
// primitive quartet, summed over lattice cells; each class reads its own
// sub-range of them.
struct TypePlan {
   int nsh[4];
   int l[4][MAXSH];
   int nl[4][MAXSH];
   int nclass;               // nsh0*nsh1*nsh2*nsh3, sa outer
   int LA0, LA, LC0, LC, L;
   int ncp;
   int cpoff[7][7];
   int nkc;                  // ket cartesian components, f in [LC0,LC]
   int kcoff[7];             // first of shell f among them
   int nbh;                  // bra Hermite triples t+u+v <= LA
   int nh;                   // Hermite triples t+u+v <= L
   int has00;                // LA0 == 0 && LC0 == 0
   int nL;                   // distinct class L values
   int Ls[16];
   int ebytes;               // contracted doubles per row: sum over classes nlabcd ncp
   int cls_code[MAXTC];      // class_code(la,lb,lc,ld) of class k
   int cls_sh[MAXTC];        // sa | sb<<4 | sc<<8 | sd<<12
   int cls_gate[MAXTC];      // index into Ls
   int cls_nl[MAXTC];        // nla*nlb*nlc*nld
   int cls_ncp[MAXTC];
   int cls_eoff[MAXTC];      // offset of class k inside a row's contracted doubles
   int co_off[4][MAXSH];     // quartet kernel: coefficients of shell s of centre t in LDS, nl x 16
   int nco_tot;
   int fun0[4][MAXSH];       // first spherical function of shell s within set t (sets list shells in order)
   int nfun[4];              // spherical functions of set t
   unsigned long long cls_mult[MAXTC];   // checksum multiplier of class k (odd)
   unsigned long long mult_all;          // sum of cls_mult
   // algorithmic FLOPs of this library's kernels (Hermite path)
   long long surv_flops;     // per slot that survives for some class: G_m and the R recursion
   long long prim_flops;     // per primitive quartet: Hermite -> cartesian transform
   long long ref_slot_flops; // per slot, summed over classes: the reference's Fm + VRR
};

// The KsGroup of class k of a set quartet stored as a KsSet: offsets of
// contraction (ia,ib,ic,id) are base + ia step[0] + ic step[1] (ac), ...
GINT_HDI KsGroup ks_from_set(const KsSet &s, const TypePlan &tp, int k)
{
   KsGroup g;
   const int sh = tp.cls_sh[k], code = tp.cls_code[k];
   const int sa = sh & 15, sb = (sh >> 4) & 15, sc = (sh >> 8) & 15, sd = (sh >> 12) & 15;
   const long long nsa = 2 * (code / 64) + 1, nsb = 2 * ((code / 16) % 4) + 1, nsc = 2 * ((code / 4) % 4) + 1,
                   nsd = 2 * (code % 4) + 1;
   const long long fa = tp.fun0[0][sa], fb = tp.fun0[1][sb], fc = tp.fun0[2][sc], fd = tp.fun0[3][sd];
   const long long ld[4] = {s.ld01 & 0xffff, s.ld01 >> 16, s.ld23 & 0xffff, s.ld23 >> 16};
   // element (row r, column c) of pair block q: rows of a (ac, ad) or b (bc, bd)
   auto pos = [&](int q, long long r, long long c) { return ((s.flags >> q) & 1) ? r * ld[q] + c : c * ld[q] + r; };
   g.base[0] = (uint32_t)(s.base[0] + pos(0, fa, fc));
   g.base[1] = (uint32_t)(s.base[1] + pos(1, fa, fd));
   g.base[2] = (uint32_t)(s.base[2] + pos(2, fb, fc));
   g.base[3] = (uint32_t)(s.base[3] + pos(3, fb, fd));
   g.step[0] = (int32_t)pos(0, nsa, 0);
   g.step[1] = (int32_t)pos(0, 0, nsc);
   g.step[2] = (int32_t)pos(1, nsa, 0);
   g.step[3] = (int32_t)pos(1, 0, nsd);
   g.step[4] = (int32_t)pos(2, nsb, 0);
   g.step[5] = (int32_t)pos(2, 0, nsc);
   g.step[6] = (int32_t)pos(3, nsb, 0);
   g.step[7] = (int32_t)pos(3, 0, nsd);
   g.ld01 = s.ld01;
   g.ld23 = s.ld23;
   g.flags = (s.flags & 0x7fu) | (uint32_t)tp.cls_nl[k] << 16;
   g.expl = 0;
   return g;
}

// Device-resident constant tables.
struct DevTables {
   const double *C0;
   int ld_C0;
   const int *grid_lo, *grid_hi;
   const double *bw;
   const double *ftable;
   const double *cell;
   const double *neighs;
   int Ng;
   int truncated;
   double R_cut;
};

// One chunk of one type's work, resident in HBM.  Primitive, row and group
// indices are absolute within the type's arrays; ac0 and eco are indexed
// from the chunk's first primitive and first row.
struct BatchDev {
   const TypePlan *tp;        // device copy
   const uint32_t *e1;        // Hermite -> cartesian stage 1 items: t u v fx fy fz (3 bits each), W slot << 18
   const uint32_t *e2;        // stage 2 items: ex ey ez (3 bits each), column << 9, block element << 16
   const RowRec *rows;
   const GroupRec *groups;
   const uint32_t *prm_row;   // row of each primitive quartet
   const uint32_t *prm_ip;    // ipa | ipb << 4 | ipc << 8 | ipd << 12
   const KsGroup *ksg;        // per class of the groups that are not KS_SETLEVEL (GroupRec::ks)
   const KsSet *kss;          // KS_SETLEVEL groups
   const KsRec *ksx;
   const double *data;
   double *ac0;               // [primitive - p0][ncp]
   uint32_t p0, p1;           // primitives of the chunk
   uint32_t r0, r1;           // rows
   uint32_t g0, g1;           // groups
   unsigned long long *stats; // STATS_SLOTS x {surviving (slot, class) pairs, checksum, K1 FLOPs} (nullable)
   uint32_t *fix;             // mixed-gate slots: (primitive, n3 | mask << 8)
   uint32_t *fix_count;       // [0] entries, [1] slots beyond fix_cap (kept by every class instead)
   uint32_t fix_cap;
   int force_fix;             // route every surviving slot through the fix-up path (test hook)
   double *fix_ws;            // fix-up workspace, ncp doubles per entry
   double *eco;               // contraction buffer: eco_row doubles per row of the chunk (k_eco)
   size_t eco_row;
   double *gmag;              // calibration: max |block element| per set quartet (nullable; absolute group index)
   const uint32_t *gflag;     // mixed precision: 1 = single-precision contraction for the set quartet (nullable)
};

// One work item of the matrix-core contraction (k_eco) of a row: class k,
// column tiles jt0 .. jt0 + 3 (16 columns each), contractions qb * 64 ..
struct EcoItem {
   uint16_t k, jt0, qb, pad;
};

struct EcoArgs {
   const EcoItem *items;      // the items of one row (the same for every row of the type)
   int n_items;
   const int *srcmaps;        // the type's class column maps (make_cq_plan)
   const int *src_off;        // [MAXTC] first entry of class k's map
   const int *eco_off;        // [MAXTC] class k's offset in a row of the contraction buffer
   size_t row_doubles;        // contraction buffer doubles per row
};

// the K layout of class k of group g
GINT_HDI KsGroup group_ks(const BatchDev &b, const TypePlan &tp, const GroupRec &g, int k)
{
   if (g.ksmode & KS_SETLEVEL) return ks_from_set(b.kss[g.ks], tp, k);
   return b.ksg[(size_t)g.ks + k];
}

struct KTargets {
   const double *P[2];
   double *K[2];
   int nspin;
   double hf_fac;
   double *out;   // integral output (KS_OUTPUT groups)
};

// position in a class's output block of integral (isa,isb,isc,isd) of
// contraction (ia,ib,ic,id): the reference's TRA order
GINT_HDI size_t tra_index(int ia, int isa, int ib, int isb, int ic, int isc, int id, int isd, int nlb, int nlc, int nld,
                          int nsa, int nsb, int nsc, int nsd)
{
   (void)nsa;
   return ((((((size_t)ia * nsa + isa) * nlb + ib) * nsb + isb) * nlc + ic) * nsc + isc) * (size_t)(nld * nsd) +
          (size_t)id * nsd + isd;
}

struct Scratch {
   double *ptr;
   size_t doubles;
   unsigned grid;   // workgroups of a k_gqs part that uses it (each its own slice)
};

constexpr unsigned SCRATCH_GRID = 2048;

// The no-allocation executor (libgint_amd_plan_execute_noalloc) takes all
// its device scratch from one static array of NOALLOC_BYTES per GPU:
//   [0, 256)                      chunk counters (fix-up entries, overflow)
//   fix-up list                   NOALLOC_FIX entries of 2 words
//   fix-up workspace              NOALLOC_FIX x (MAX_NCP + 16) doubles
//   gqs workspace                 NOALLOC_GWS_BYTES
//   primitive blocks (ac0)        the rest, NOALLOC_AC0_BYTES
constexpr size_t NOALLOC_BYTES = 176ull << 20;
constexpr int MAX_NCP = 84 * 84;              // (e0|f0) doubles of an (f f|f f) type
constexpr size_t NOALLOC_FIX = 256;
constexpr size_t NOALLOC_OFF_FIX = 256;
constexpr size_t NOALLOC_OFF_FIXWS = NOALLOC_OFF_FIX + ((2 * 4 * NOALLOC_FIX + 255) & ~(size_t)255);
constexpr size_t NOALLOC_OFF_GWS = NOALLOC_OFF_FIXWS + ((NOALLOC_FIX * (MAX_NCP + 16) * 8 + 255) & ~(size_t)255);
constexpr size_t NOALLOC_GWS_BYTES = 32ull << 20;
constexpr size_t NOALLOC_OFF_AC0 = NOALLOC_OFF_GWS + NOALLOC_GWS_BYTES;
constexpr size_t NOALLOC_AC0_BYTES = NOALLOC_BYTES - NOALLOC_OFF_AC0;
static_assert(NOALLOC_AC0_BYTES >= (120ull << 20), "no-allocation scratch layout");
constexpr unsigned STATS_SLOTS = 4096;   // counter triples the kernels spread their statistics over

// LDS plan of one class in the quartet kernel
struct CqLayout {
   int lb;          // contractions per pass
   int lws;         // workspace doubles per contraction: 3 ycap + Nc
   int o_ws;        // workspace in LDS (lb x lws), unless ws_global
   int o_kp;        // coefficient products [lb][CQ_PCH]
   int o_I;         // spherical integrals of every contraction [nq][Ns] (reduce)
   int o_src;       // first entry of the class's block map (class element -> type element) in the type's maps
   int reduce;
   int ws_global;   // workspace in HBM (gws + blockIdx * lb * lws)
   int skip;        // the class runs in its own templated kernel instead
   int part;        // which of the GQS_PARTS concurrent launches takes the class
   int bytes;
   int eco;         // contracted on the matrix cores (k_eco): k_gqs reads the contraction buffer
   int eco_off;     // the class's offset in a row of the contraction buffer
   int o_kron;      // LDS of the Kronecker path: bra components x ket spherical pairs (whole 16x16 tiles), T tables
   int kron_ws;     // doubles of the Kronecker path's row maps (up to 8 rows); 0: the class cannot take the path
   int o_kmap;      // LDS offset of those maps, or -1: in the workgroup's HBM slice
   int kron;        // the class takes the Kronecker path (a schedule of the autotuner)
   int kron_kb;     // its contractions per pass
   int o_kc;        // LDS of one contraction's blocks and CQ_PCH coefficient products for the Kronecker path
                    // with the class's own contraction (no eco), or -1: that schedule does not fit
   int o_kacc;      // LDS of the K accumulators (the four K blocks of a set quartet, summed over its
                    // contractions before the atomics), or -1
   int kacc;        // doubles of the accumulators per spin
   int kacc_cap;    // doubles reserved for them (kacc or 2 kacc: one or two spins)
};

// address of the no-allocation scratch array on the current device, looked
// up (hipGetSymbolAddress) once per device; null if the lookup failed
void *noalloc_scratch();

// launchers (kernels.hip), asynchronous on `stream`; 0 or a negative code
// zeroes n words (a kernel, not a memset node, so that the launch sequence
// can be captured in a graph)
int launch_zero_u32(uint32_t *p, int n, void *stream);
int launch_fm_r(const TypePlan &tp, const DevTables &t, const BatchDev &b, void *stream);
int launch_fixup(const TypePlan &tp, const DevTables &t, const BatchDev &b, uint32_t n_fix, void *stream);
// one-wavefront-per-set-quartet kernels of the classes with at most 128
// block columns (kernels_tq.hip)
bool tq_has(int code);
int launch_tq(int cls_k, const TypePlan &tp, const DevTables &t, const BatchDev &b, const KTargets &kt,
              const int *srcmap, void *stream);
// per-class LDS plans and block maps of a type for the quartet kernel
void make_cq_plan(const TypePlan &tp, const std::vector<ClassPlan> &cps, std::vector<CqLayout> &lays,
                  std::vector<int> &srcmaps, int *lds_bytes, int *gws_stride);
// the k_gqs classes of a type over its GQS_PARTS launches, balanced by cost
void assign_parts(const TypePlan &tp, const std::vector<ClassPlan> &cps, std::vector<CqLayout> &lays);
// the classes of a type that k_gqs takes are split into this many launches,
// balanced by their estimated cost, which run concurrently on separate streams
constexpr int GQS_PARTS = 4;
// npc: the classes of the type that this part takes (work items are
// (set quartet, class) pairs)
constexpr unsigned GQS_MAX_GRID = 8192;
int launch_gqs(const TypePlan &tp, const ClassPlan *d_cps, const CqLayout *d_lays, const int *d_srcmaps, int lds_bytes,
               int gws_stride, int part, int npc, const DevTables &t, const BatchDev &b, const KTargets &kt, Scratch s,
               void *stream, bool wide = false);
// contraction of primitive blocks on the matrix cores for the classes with
// CqLayout::eco (kernels_eco.hip): k_gqs classes with at least this many
// contractions
constexpr int ECO_MIN_NQ = 8;
// classes with at least this many contractions have a place in the
// contraction buffer (the autotuner may move them to the matrix cores)
constexpr int ECO_CAND_NQ = 4;
int launch_eco(const TypePlan &tp, const BatchDev &b, const EcoArgs &e, void *stream);
// mixed precision: max |P| (*out, zeroed first) and the route of each set
// quartet of a chunk from its calibrated magnitude
int launch_absmax(const double *P, size_t n, unsigned long long *out, bool zero, void *stream);
int launch_mixed_flags(const TypePlan &tp, const BatchDev &b, const double *gmag, const unsigned long long *pmax,
                       double tau, uint32_t *gflag, unsigned long long *count, void *stream);
// D (16 x 16, row-major) = A (16 x 4) B (4 x 16) through the matrix-core
// tile the contraction uses, with its lane layout; 0 or -1
int mfma_selftest(const double *A, const double *B, double *D, int prec);
// LDS bytes of the fused kernel for a type (0 if it does not fit)
size_t fm_r_lds_bytes(const TypePlan &tp);

ClassPlan make_plan(int la, int lb, int lc, int ld);
// fills everything from nsh, l and nl
void make_type_plan(TypePlan &tp);
// item descriptors of the Hermite -> cartesian transform of a type
void make_type_items(const TypePlan &tp, std::vector<uint32_t> &e1, std::vector<uint32_t> &e2);
uint64_t class_mult(int code);
void boys_table(double *ftable);
int device_math_selftest(const double *x, const double *y, int n, double *out);

}  // namespace gint
