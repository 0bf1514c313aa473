// Interface between the host engine (engine.cpp, built with g++ -fopenmp)
// and the gfx950 kernels (kernels.hip, built with hipcc): record layouts in
// HBM, the per-class plan passed by value to the kernels, and the launchers.
// Plain types only.
#pragma once
#include <stdint.h>
#include <stddef.h>

namespace gint {

constexpr int MAXL = 3;          // highest angular momentum of a shell (f)
constexpr int MAXLL = 4 * MAXL;  // highest total angular momentum of a class
constexpr int MAXNL = 4;         // contractions per shell and centre
constexpr int MAXNP = 16;        // primitives per set
constexpr int MAXNG = 256;       // lattice vectors
constexpr int NCLASS = 256;      // (la,lb,lc,ld), each 0..3

inline constexpr int nco(int l) { return (l + 1) * (l + 2) / 2; }
inline constexpr int nsph(int l) { return 2 * l + 1; }
inline constexpr int class_index(int la, int lb, int lc, int ld) { return ((la * 4 + lb) * 4 + lc) * 4 + ld; }

// One row: a set quartet (i,j,k,l) at bra/ket images (n1,n2) for one
// angular-momentum class, with its primitive quartets.  All rows of one
// set quartet and class share q_off (their integrals are summed).
struct RowRec {
   uint32_t prm_off;   // first primitive of the row in the batch
   uint32_t n_prm;
   uint32_t q_off;     // first contracted quartet of the row in the batch
   uint32_t nlnp;      // nl[0..3] in 4-bit fields, then np[0..3]-1 in 4-bit fields
   uint32_t n12;       // n1 | n2 << 16
   uint32_t R[4];      // offsets of the four centres in the basis data
   uint32_t Z[4];      // offsets of the exponents
   uint32_t K[4];      // offsets of the normalised coefficients (nl x np)
   uint32_t e_off;     // first contraction slot of the row in the contracted-block array
   uint32_t set_ij;    // set indices i | j << 16 (survivor checksum only)
   uint32_t set_kl;    // set indices k | l << 16
};
static_assert(sizeof(RowRec) == 80, "RowRec layout");

// One contracted quartet (one add_qrtt call): where its exchange terms go.
struct KsRec {
   uint32_t off[4];    // ac, ad, bc, bd block offsets into P / K
   uint32_t ld01;      // ld_ac | ld_ad << 16
   uint32_t ld23;      // ld_bc | ld_bd << 16
   uint32_t flags;     // bit t: transpose flag of block t; bits 4..6: symmetry code
   uint32_t row0;      // first row of its set quartet
   uint32_t row1;      // one past the last row
   uint32_t l;         // contraction combination index inside the quartet
   uint32_t pad[2];
};
static_assert(sizeof(KsRec) == 48, "KsRec layout");

// Everything a kernel needs to know about one class; passed by value.
struct ClassPlan {
   int la, lb, lc, ld;
   int L, LA, LC;            // L = la+lb+lc+ld, LA = la+lb, LC = lc+ld
   int ncp;                  // size of the (e0|f0) block set, e in [la,LA], f in [lc,LC]
   int cpoff[7][7];          // offset of block (e,f) inside the set
   int vws;                  // VRR workspace doubles per slot
   int voff[7][7];           // offset of VRR block (a,c) ([m][ia][ic], m = 0..L-a-c)
   int amin[7];              // lowest a needed at ket level c
   int Nc, Ns;               // cartesian / spherical sizes of the class
   // HRR layers.  Ket layer dd (0..ld) holds blocks (e, c) for e in [la,LA],
   // c in [lc, LC-dd], layout [ie][ic][id], e outer; layer 0 is the (e0|f0)
   // set itself.  Bra layer bb (0..lb) holds blocks a in [la, LA-bb], layout
   // [ia][ib][icd]; bra layer 0 is the last ket layer.
   int koff[4][7][7];        // offset of ket block (e,c) in layer dd
   int ksize[4];
   int boff[4][7];           // offset of bra block a in layer bb
   int bsize[4];
   int ycap;                 // capacity of one layer buffer (>= ncp, Nc, every layer)
   int hrr_ws;               // HRR workspace doubles per quartet: 3 ycap + Nc
   // algorithmic FLOPs fixed by the reference recurrence chain (not by
   // this implementation): per surviving slot, per contracted quartet
   int fm_flops;             // 52 + 211 (L+1)
   int vrr_flops;            // reference plan VRR ops, weights 3/8/3/8/11/6 per element
   int hrr_flops;            // 2 per HRR element
   int sph_flops;            // 2 m n k over the four stages
};

// Device-resident, per-process constant tables.
struct DevTables {
   const double *C0;        // patch coefficients, ld_C0 per patch
   int ld_C0;
   const int *grid_lo, *grid_hi;
   const double *bw;        // 207 x 4
   const double *ftable;    // 121 x 28 Boys table
   const double *cell;      // 18
   const double *neighs;    // 3 x Ng
   int Ng;
   int truncated;
   double R_cut;
};

// One class batch resident in HBM.
struct BatchDev {
   const RowRec *rows;
   const uint32_t *prm_row;   // row of each primitive
   const uint32_t *prm_ip;    // ipa | ipb << 4 | ipc << 8 | ipd << 12
   const KsRec *ks;
   const double *data;        // basis data (positions, exponents, coefficients)
   double *ac0;               // [n_prm][ncp]
   double *eco;               // per row [nlabcd][ncp]
   uint32_t n_rows, n_prm, n_q;
   unsigned long long *stats; // [0] surviving slots, [1] slot checksum, [2] fm+vrr FLOPs (may be null)
};

struct KTargets {
   const double *P[2];
   double *K[2];
   int nspin;
   double hf_fac;
};

// Scratch in HBM for classes whose per-wavefront workspace exceeds the LDS
// budget; owned by the caller (one per stream).
struct Scratch {
   double *ptr;
   size_t doubles;
};

// launchers (kernels.hip); all asynchronous on `stream`, return 0 or a
// negative code (-1 launch failure, -2 scratch too small)
int launch_fm_vrr(const ClassPlan &cp, const DevTables &t, const BatchDev &b, Scratch s, void *stream);
// contraction of rows [row_begin,row_end) over the primitives [prm_begin,prm_end)
// whose (e0|f0) sums sit in b.ac0 (indexed from prm_begin); accumulates
int launch_eco_range(const ClassPlan &cp, const BatchDev &b, uint32_t row_begin, uint32_t row_end,
                     uint32_t prm_begin, uint32_t prm_end, void *stream);
int launch_hrr_sph_ks(const ClassPlan &cp, const DevTables &t, const BatchDev &b, const KTargets &kt, Scratch s,
                      void *stream);
// workgroups that get an HBM workspace slice when a class does not fit in LDS
constexpr unsigned SCRATCH_GRID = 2048;
// HBM scratch doubles the class needs (0 if its workspaces fit in LDS)
size_t scratch_doubles(const ClassPlan &cp);
// fill a ClassPlan (plan.cpp)
ClassPlan make_plan(int la, int lb, int lc, int ld);
// 121 x 28 Boys table (boys_table.cpp)
void boys_table(double *ftable);
// device self-test of the scalar functions that decide screening:
// out[4i..4i+3] = {gint_exp(x), sqrt(|x|), 1/x, x*y/(x+y)} evaluated on the GPU
int device_math_selftest(const double *x, const double *y, int n, double *out);

}  // namespace gint
