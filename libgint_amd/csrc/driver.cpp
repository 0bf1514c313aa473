// Caller-side driver of the libgint_* ABI: the part of CP2K's HFX module
// that a user of this library keeps (it is not in the reference repository).
// It builds deterministic synthetic systems, lays out P and K the way CP2K
// passes them (one column-major block per atom pair i <= j, leading
// dimension = functions of atom i), screens quartets and drives the call
// protocol of reference src/interface_libgint.F90:404-462.
//
//  * systems: water-like molecules (O, H at 0.96 A and 104.5 deg) and
//    optional transition-metal-like centres at seeded positions, one basis
//    set per atom, exponents log-uniform in [0.1, 5e3] bohr^-2,
//    contraction coefficients N(0,1)
//  * density: P = (C C^T) o exp(-r_ij / 2 A), C of shape nbf x 40, N(0,1)
//  * symmetry: atom-level 8-fold (i <= j, k <= l, (k,l) <= (i,j)) with
//    complete image sums (n1 over b, n2 over d, n3 inside the library), so
//    a quartet weighs 2 / 2^([i=j] + [k=l] + [(i,j)=(k,l)]) and the final
//    K is the symmetric part of what the library accumulates (the
//    reference's own harness does the same, src/main.cpp:271-293,492-547)
//  * screening: primitive-pair Schwarz estimate x max|P| over the four
//    exchange blocks >= eps, and the exact truncation test on the nearest
//    ket image (G_m = 0 when T > R^2 + 11 R + 50)
//
// A recorded stream packs one 32-bit word per primitive quartet; the
// add_qrt / add_qrtt calls of a set quartet are regenerated at replay.
#include <omp.h>
#include <stdint.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gint_math.h"

namespace {

constexpr double BOHR_PER_ANG = 1.8897261254578281;

struct Rng {
   uint64_t s[4];
   explicit Rng(uint64_t seed)
   {
      uint64_t x = seed;
      for (auto &w : s) {
         x += 0x9E3779B97F4A7C15ULL;  // splitmix64
         uint64_t z = x;
         z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
         z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
         w = z ^ (z >> 31);
      }
   }
   static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
   uint64_t next()
   {  // xoshiro256**
      const uint64_t r = rotl(s[1] * 5, 7) * 9;
      const uint64_t t = s[1] << 17;
      s[2] ^= s[0];
      s[3] ^= s[1];
      s[1] ^= s[2];
      s[0] ^= s[3];
      s[2] ^= t;
      s[3] = rotl(s[3], 45);
      return r;
   }
   double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
   int range(int lo, int hi) { return std::min(hi, lo + (int)(uniform() * (hi - lo + 1))); }
   double normal()
   {  // Box-Muller
      double u1 = uniform(), u2 = uniform();
      if (u1 < 1e-300) u1 = 1e-300;
      return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
   }
};

struct Shell {
   int l, nl;
   std::vector<double> K;   // nl x np as passed to set_Atom_L
   double cmax;             // max |normalised coefficient| over contractions and primitives
   std::vector<double> cmax_p;
};
struct Basis {
   int np = 0;
   double Z[16];
   std::vector<Shell> sh;
   int nfun = 0;
   std::vector<double> cmax_p;  // per primitive, over shells
};
struct Atom {
   double R[3];
   int kind;
};

}  // namespace

// parameters of a synthetic system (mirrored by libgint_amd/driver.py)
struct GdrvParams {
   int n_water;
   int n_tm;
   double box_A;
   int image_shells;   // 0: 1 cell, 1: 27 ({-1,0,1}^3), 2: 125 ({-2..2}^3), 3: the 256 shortest lattice vectors
   int basis;          // 0: O 3s2p1d / H 2s1p; 1: O 2s1p / H 1s; 4: O 4s3p2d1f / H 3s2p1d
   int np_min, np_max;
   uint64_t seed;
   double rcut_A;
   double eps;
   double min_oo_A;
   int symmetric;      // 1: unique quartets with weights; 0: every ordered quartet, weight 1/4
   double cell_A;      // cell edge; 0 -> box_A (a larger cell makes the minimum image the identity)
   int np_tm;          // primitives of the spdf centres (16 in the preset workloads; fewer for tests)
   int max_cells;      // > 0: keep only the first max_cells lattice vectors (nearest first)
   int ks_mix;         // 1: vary the order of the add_qrtt calls so that K layouts of every kind occur
   double zmin, zmax;  // range of the log-uniform exponent draw (bohr^-2)
   int shift_atom;     // >= 0: this atom is moved by shift_cells lattice vectors after placement
   int shift_cells[3];
   int replicate;      // > 1: the system repeated this many times along x in a cell that many times longer,
                       // with the density of every copy pair equal to the original pair's
};

struct GdrvSystem {
   GdrvParams p;
   std::vector<Atom> atoms;
   std::vector<Basis> kinds;  // 0 O, 1 H, 2 TM
   double cell18[18];
   std::vector<double> neighs;
   int nbf = 0;
   std::vector<int> fun0;
   std::vector<uint32_t> blk;  // natom x natom, valid for i <= j
   size_t P_size = 0;
   std::vector<double> P;
   std::vector<double> Pmax;   // natom x natom
   double R_cut = 0;
};

namespace {

double norm_psi(const double *K, const double *z, int np, int l)
{  // reference src/extern_functions.cpp:95-110
   double ans = 0.0;
   for (int i = 0; i < np; i++)
      for (int j = 0; j < np; j++) ans += K[i] * K[j] / std::pow(z[i] + z[j], 1.5 + l);
   ans *= std::pow(M_PI, 1.5) / std::pow(2.0, l);
   return std::sqrt(ans);
}

Basis make_basis(Rng &rng, const std::vector<std::pair<int, int>> &lnl, int np, double zmin, double zmax)
{
   Basis b;
   b.np = np;
   std::vector<double> z(np);
   for (int i = 0; i < np; i++) z[i] = std::exp(std::log(zmin) + rng.uniform() * (std::log(zmax) - std::log(zmin)));
   std::sort(z.begin(), z.end(), [](double a, double c) { return a > c; });
   for (int i = 0; i < np; i++) b.Z[i] = z[i];
   b.cmax_p.assign(np, 0.0);
   for (auto &q : lnl) {
      Shell s;
      s.l = q.first;
      s.nl = q.second;
      s.K.resize((size_t)s.nl * np);
      for (auto &v : s.K) v = rng.normal();
      s.cmax = 0;
      s.cmax_p.assign(np, 0.0);
      for (int n = 0; n < s.nl; n++) {
         const double nrm = norm_psi(&s.K[(size_t)n * np], b.Z, np, s.l);
         for (int i = 0; i < np; i++) {
            const double c = std::fabs(s.K[(size_t)n * np + i] / nrm);
            s.cmax = std::max(s.cmax, c);
            s.cmax_p[i] = std::max(s.cmax_p[i], c);
            b.cmax_p[i] = std::max(b.cmax_p[i], c);
         }
      }
      b.nfun += s.nl * (2 * s.l + 1);
      b.sh.push_back(s);
   }
   return b;
}

// lattice vectors of the {-r..r}^3 block of a rectangular cell of edges
// e[0..2], nearest first by cell indices (stable: ties keep the loop order);
// r = 3 keeps the 256 shortest, the reference's cell limit (src/define.h:42),
// which covers |g|^2 <= 14 (251 vectors) and 5 of the 24 at |g|^2 = 17
void make_neighs(GdrvSystem &S, const double e[3])
{
   const GdrvParams &p = S.p;
   std::vector<std::array<int, 3>> v;
   const int r = p.image_shells <= 0 ? 0 : std::min(p.image_shells, 3);
   for (int a = -r; a <= r; a++)
      for (int b = -r; b <= r; b++)
         for (int c = -r; c <= r; c++) v.push_back({a, b, c});
   std::stable_sort(v.begin(), v.end(), [](const std::array<int, 3> &x, const std::array<int, 3> &y) {
      return x[0] * x[0] + x[1] * x[1] + x[2] * x[2] < y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
   });
   if (p.image_shells >= 3 && v.size() > 256) v.resize(256);
   if (p.max_cells > 0 && (size_t)p.max_cells < v.size()) v.resize(p.max_cells);
   S.neighs.clear();
   for (auto &g : v)
      for (int k = 0; k < 3; k++) S.neighs.push_back(g[k] * e[k]);
}

// function offsets, pair blocks (i <= j, column-major, ld = functions of i)
// and the P size of the current atom list
void layout_blocks(GdrvSystem &S)
{
   const int na = (int)S.atoms.size();
   S.fun0.resize(na);
   S.nbf = 0;
   for (int a = 0; a < na; a++) {
      S.fun0[a] = S.nbf;
      S.nbf += S.kinds[S.atoms[a].kind].nfun;
   }
   S.blk.assign((size_t)na * na, 0);
   size_t off = 0;
   for (int i = 0; i < na; i++)
      for (int j = i; j < na; j++) {
         S.blk[(size_t)i * na + j] = (uint32_t)off;
         off += (size_t)S.kinds[S.atoms[i].kind].nfun * S.kinds[S.atoms[j].kind].nfun;
      }
   S.P_size = off;
}

void make_pmax(GdrvSystem &S)
{
   const int na = (int)S.atoms.size();
   S.Pmax.assign((size_t)na * na, 0.0);
   for (int a = 0; a < na; a++)
      for (int b = a; b < na; b++) {
         const size_t base = S.blk[(size_t)a * na + b];
         const size_t n = (size_t)S.kinds[S.atoms[a].kind].nfun * S.kinds[S.atoms[b].kind].nfun;
         double m = 0;
         for (size_t t = 0; t < n; t++) m = std::max(m, std::fabs(S.P[base + t]));
         S.Pmax[(size_t)a * na + b] = S.Pmax[(size_t)b * na + a] = m;
      }
}

// the system repeated nx times along x: copy k of atom a is atom k na + a,
// at R + k edge x; the cell is nx edges long in x; the density block of two
// copies is the original block of the two atoms (so that, summed over the
// copies of its second atom, the exchange matrix of the longer cell is the
// original one)
void replicate_x(GdrvSystem &S, int nx, double edge)
{
   const int na = (int)S.atoms.size();
   const std::vector<Atom> a0 = S.atoms;
   const std::vector<uint32_t> blk0 = S.blk;
   const std::vector<double> P0 = S.P;
   S.atoms.clear();
   for (int k = 0; k < nx; k++)
      for (int a = 0; a < na; a++) {
         Atom t = a0[a];
         t.R[0] += k * edge;
         S.atoms.push_back(t);
      }
   S.cell18[0] = nx * edge;
   S.cell18[9] = 1.0 / (nx * edge);
   const double e[3] = {nx * edge, S.cell18[4], S.cell18[8]};
   make_neighs(S, e);
   layout_blocks(S);
   const int nn = (int)S.atoms.size();
   S.P.assign(S.P_size, 0.0);
   for (int I = 0; I < nn; I++)
      for (int J = I; J < nn; J++) {
         const int i = I % na, j = J % na;
         const int ni = S.kinds[a0[i].kind].nfun, nj = S.kinds[a0[j].kind].nfun;
         double *B = &S.P[S.blk[(size_t)I * nn + J]];
         for (int r = 0; r < ni; r++)
            for (int c = 0; c < nj; c++)
               B[r + (size_t)c * ni] = i <= j ? P0[blk0[(size_t)i * na + j] + r + (size_t)c * ni]
                                              : P0[blk0[(size_t)j * na + i] + c + (size_t)r * nj];
      }
   make_pmax(S);
}

void build(GdrvSystem &S)
{
   const GdrvParams &p = S.p;
   Rng rng(p.seed);
   const double zmin = p.zmin > 0 ? p.zmin : 0.1, zmax = p.zmax > 0 ? p.zmax : 5e3;
   const double box = p.box_A * BOHR_PER_ANG;
   const double cedge = (p.cell_A > 0 ? p.cell_A : p.box_A) * BOHR_PER_ANG;
   for (int i = 0; i < 9; i++) {
      S.cell18[i] = (i % 4 == 0) ? cedge : 0.0;
      S.cell18[9 + i] = (i % 4 == 0) ? 1.0 / cedge : 0.0;
   }
   {
      const double e[3] = {cedge, cedge, cedge};
      make_neighs(S, e);
   }
   if (p.basis == 4) {
      S.kinds.push_back(make_basis(rng, {{0, 4}, {1, 3}, {2, 2}, {3, 1}}, rng.range(p.np_min, p.np_max), zmin, zmax));
      S.kinds.push_back(make_basis(rng, {{0, 3}, {1, 2}, {2, 1}}, rng.range(p.np_min, p.np_max), zmin, zmax));
   } else if (p.basis == 1) {
      S.kinds.push_back(make_basis(rng, {{0, 2}, {1, 1}}, rng.range(p.np_min, p.np_max), zmin, zmax));
      S.kinds.push_back(make_basis(rng, {{0, 1}}, rng.range(p.np_min, p.np_max), zmin, zmax));
   } else {
      S.kinds.push_back(make_basis(rng, {{0, 3}, {1, 2}, {2, 1}}, rng.range(p.np_min, p.np_max), zmin, zmax));
      S.kinds.push_back(make_basis(rng, {{0, 2}, {1, 1}}, rng.range(p.np_min, p.np_max), zmin, zmax));
   }
   S.kinds.push_back(make_basis(rng, {{0, 2}, {1, 2}, {2, 2}, {3, 4}}, p.np_tm > 0 ? p.np_tm : 16, zmin, zmax));

   const double dOH = 0.96 * BOHR_PER_ANG, ang = 104.5 * M_PI / 180.0;
   const double dmin = p.min_oo_A * BOHR_PER_ANG;
   std::vector<std::array<double, 3>> heavy;
   auto place = [&](double *R) {
      for (int tries = 0; tries < 100000; tries++) {
         for (int k = 0; k < 3; k++) R[k] = rng.uniform() * box;
         bool ok = true;
         for (auto &o : heavy) {
            double m[3];
            gint_pbc(o.data(), R, S.cell18, m);
            if (std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) < dmin) {
               ok = false;
               break;
            }
         }
         if (ok) return;
      }
   };
   for (int w = 0; w < p.n_water; w++) {
      Atom O;
      O.kind = 0;
      place(O.R);
      heavy.push_back({O.R[0], O.R[1], O.R[2]});
      double u[3] = {rng.normal(), rng.normal(), rng.normal()};
      const double nu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      for (double &x : u) x /= nu;
      double t[3] = {rng.normal(), rng.normal(), rng.normal()};
      const double dt = t[0] * u[0] + t[1] * u[1] + t[2] * u[2];
      for (int k = 0; k < 3; k++) t[k] -= dt * u[k];
      const double nt = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      for (double &x : t) x /= nt;
      S.atoms.push_back(O);
      for (int h = 0; h < 2; h++) {
         const double th = (h == 0 ? 0.5 : -0.5) * ang;
         Atom H;
         H.kind = 1;
         for (int k = 0; k < 3; k++) H.R[k] = O.R[k] + dOH * (std::cos(th) * u[k] + std::sin(th) * t[k]);
         S.atoms.push_back(H);
      }
   }
   for (int m = 0; m < p.n_tm; m++) {
      Atom T;
      T.kind = 2;
      place(T.R);
      heavy.push_back({T.R[0], T.R[1], T.R[2]});
      S.atoms.push_back(T);
   }

   if (p.shift_atom >= 0 && p.shift_atom < (int)S.atoms.size())
      for (int k = 0; k < 3; k++) S.atoms[p.shift_atom].R[k] += p.shift_cells[k] * cedge;

   const int na = (int)S.atoms.size();
   layout_blocks(S);

   const int nocc = 40, nbf = S.nbf;
   std::vector<double> C((size_t)nbf * nocc);
   for (auto &x : C) x = rng.normal();
   std::vector<int> atom_of(nbf);
   for (int a = 0; a < na; a++)
      for (int f = 0; f < S.kinds[S.atoms[a].kind].nfun; f++) atom_of[S.fun0[a] + f] = a;
   std::vector<double> dist((size_t)na * na);
   for (int a = 0; a < na; a++)
      for (int b = 0; b < na; b++) {
         double m[3];
         gint_pbc(S.atoms[a].R, S.atoms[b].R, S.cell18, m);
         dist[(size_t)a * na + b] = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      }
   S.P.assign(S.P_size, 0.0);
   const double decay = 2.0 * BOHR_PER_ANG;
#pragma omp parallel for schedule(dynamic, 4)
   for (int mu = 0; mu < nbf; mu++)
      for (int nu = mu; nu < nbf; nu++) {
         double s = 0.0;
         for (int o = 0; o < nocc; o++) s += C[(size_t)mu * nocc + o] * C[(size_t)nu * nocc + o];
         const int a = atom_of[mu], b = atom_of[nu];
         const double v = s * std::exp(-dist[(size_t)a * na + b] / decay);
         const int lda = S.kinds[S.atoms[a].kind].nfun;
         const size_t base = S.blk[(size_t)a * na + b];
         const int ra = mu - S.fun0[a], cb = nu - S.fun0[b];
         S.P[base + ra + (size_t)cb * lda] = v;
         if (a == b) S.P[base + cb + (size_t)ra * lda] = v;
      }
   make_pmax(S);
   if (p.replicate > 1) replicate_x(S, p.replicate, cedge);
   S.R_cut = p.rcut_A * BOHR_PER_ANG;
}

// ---------------------------------------------------------------- pairs
struct PrimPair {
   uint8_t ia, ib;
   double Q, zeta;
   double P[3];
};
struct PairImage {
   int n1;
   double Qmax;
   std::vector<PrimPair> pp;  // sorted by Q, descending
};

// Schwarz-type estimate of a primitive pair: sqrt of its s-type (ab|ab)
// Coulomb integral, times the largest normalised coefficients
std::vector<PairImage> pair_images(const GdrvSystem &S, int i, int j, double qmin)
{
   std::vector<PairImage> out;
   const Basis &A = S.kinds[S.atoms[i].kind], &B = S.kinds[S.atoms[j].kind];
   const double *RA = S.atoms[i].R;
   double ABs[3];
   gint_pbc(RA, S.atoms[j].R, S.cell18, ABs);
   const int Ng = (int)S.neighs.size() / 3;
   for (int n1 = 0; n1 < Ng; n1++) {
      double RB[3];
      for (int k = 0; k < 3; k++) RB[k] = RA[k] + ABs[k] + S.neighs[3 * n1 + k];
      const double R2 = (RA[0] - RB[0]) * (RA[0] - RB[0]) + (RA[1] - RB[1]) * (RA[1] - RB[1]) + (RA[2] - RB[2]) * (RA[2] - RB[2]);
      PairImage pi;
      pi.n1 = n1;
      pi.Qmax = 0;
      for (int a = 0; a < A.np; a++)
         for (int b = 0; b < B.np; b++) {
            const double za = A.Z[a], zb = B.Z[b], z = za + zb;
            const double Kab = std::exp(-za * zb / z * R2);
            const double Q = A.cmax_p[a] * B.cmax_p[b] * Kab * std::sqrt(2.0 * std::pow(M_PI, 2.5) / (z * z * std::sqrt(2.0 * z)));
            if (Q < qmin) continue;
            PrimPair pp;
            pp.ia = (uint8_t)a;
            pp.ib = (uint8_t)b;
            pp.Q = Q;
            pp.zeta = z;
            for (int k = 0; k < 3; k++) pp.P[k] = (za * RA[k] + zb * RB[k]) / z;
            pi.pp.push_back(pp);
            pi.Qmax = std::max(pi.Qmax, Q);
         }
      if (!pi.pp.empty()) {
         std::sort(pi.pp.begin(), pi.pp.end(), [](const PrimPair &x, const PrimPair &y) { return x.Q > y.Q; });
         out.push_back(std::move(pi));
      }
   }
   return out;
}

inline uint32_t pack16(int a, int b) { return (uint32_t)(a & 0xffff) | (uint32_t)(b & 0xffff) << 16; }

}  // namespace

// ---------------------------------------------------------------- stream
struct GdrvStream {
   std::vector<uint32_t> w;
   int64_t n_sets = 0, n_rows = 0, n_prm = 0, n_qrtt = 0;
   std::vector<int64_t> set_word;  // first word of every set
};

typedef struct {
   void (*add_prm)(void *, int, int, int, int);
   void (*add_shell)(void *, int, int, int, int, int, int);
   void (*add_qrt)(void *, int, int, int, int, int, int, int, int);
   void (*add_qrtt)(void *, double, int, int, int, int, int, int, int, int, int, int, int, int, unsigned, unsigned,
                    unsigned, unsigned, bool, bool, bool, bool);
   void (*add_set)(void *);
   void (*set_Atom)(void *, int, double *, double *, int);
   void (*set_Atom_L)(void *, int, int, int, double *);
   // optional: all add_qrt / add_qrtt calls of a set quartet at once
   // (libgint_amd_add_set_blocks); null when the library has none
   void (*add_set_blocks)(void *, double, int, int, int, int, unsigned, unsigned, unsigned, unsigned, bool, bool,
                          bool, bool);
   // optional: integral output of one class (libgint_amd_add_tra)
   void (*add_tra)(void *, int, int, int, int, uint64_t);
} GdrvAbi;

namespace {

constexpr uint32_t OP_PRM = 0u, OP_SHELL = 1u, OP_SET = 2u;

// a set quartet (i,j,k,l) belongs to the sample of 1 in `sample` when this
// hash of its indices is a multiple of `sample` (independent of the order
// in which the stream is recorded)
inline bool in_sample(int i, int j, int k, int l, int64_t sample)
{
   if (sample <= 1) return true;
   uint64_t z = ((uint64_t)(uint32_t)pack16(i, j) << 32 | pack16(k, l)) + 0x9E3779B97F4A7C15ULL;
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   z ^= z >> 31;
   return z % (uint64_t)sample == 0;
}

void record_bra(const GdrvSystem &S, const std::vector<std::vector<PairImage>> &pairs,
                const std::vector<std::pair<int, int>> &plist, size_t ib, std::vector<uint32_t> &out,
                int64_t cnt[4], int64_t sample)
{
   const int na = (int)S.atoms.size();
   const double eps = S.p.eps;
   const int i = plist[ib].first, j = plist[ib].second;
   const std::vector<PairImage> &BI = pairs[ib];
   if (BI.empty()) return;
   double qbra = 0;
   for (auto &x : BI) qbra = std::max(qbra, x.Qmax);
   const size_t kend = S.p.symmetric ? ib + 1 : plist.size();
   std::vector<uint32_t> prm;
   for (size_t kb = 0; kb < kend; kb++) {
      const int k = plist[kb].first, l = plist[kb].second;
      const std::vector<PairImage> &KI = pairs[kb];
      if (KI.empty() || !in_sample(i, j, k, l, sample)) continue;
      const double Pm = std::max(std::max(S.Pmax[(size_t)i * na + k], S.Pmax[(size_t)i * na + l]),
                                 std::max(S.Pmax[(size_t)j * na + k], S.Pmax[(size_t)j * na + l]));
      double qket = 0;
      for (auto &x : KI) qket = std::max(qket, x.Qmax);
      if (qbra * qket * Pm < eps) continue;
      bool any = false;
      for (const PairImage &b1 : BI) {
         if (b1.Qmax * qket * Pm < eps) continue;
         for (const PairImage &k2 : KI) {
            if (b1.Qmax * k2.Qmax * Pm < eps) continue;
            prm.clear();
            for (const PrimPair &p1 : b1.pp) {
               if (p1.Q * k2.Qmax * Pm < eps) break;
               for (const PrimPair &p2 : k2.pp) {
                  if (p1.Q * p2.Q * Pm < eps) break;
                  // exact truncation test on the nearest ket image
                  const double rho = p1.zeta * p2.zeta / (p1.zeta + p2.zeta);
                  double d[3];
                  gint_pbc(p2.P, p1.P, S.cell18, d);
                  const double T = rho * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                  const double R = S.R_cut * std::sqrt(rho);
                  if (T > R * R + 11.0 * R + 50.0) continue;
                  prm.push_back((uint32_t)p1.ia | (uint32_t)p1.ib << 4 | (uint32_t)p2.ia << 8 | (uint32_t)p2.ib << 12);
               }
            }
            if (prm.empty()) continue;
            out.insert(out.end(), prm.begin(), prm.end());
            out.push_back(OP_SHELL << 30 | (uint32_t)b1.n1 | (uint32_t)k2.n1 << 8);
            out.push_back(pack16(i, j));
            out.push_back(pack16(k, l));
            cnt[1]++;
            cnt[2] += (int64_t)prm.size();
            any = true;
         }
      }
      if (any) {
         int code;
         if (S.p.symmetric) {
            int e = (i == j) + (k == l) + (i == k && j == l);
            code = e;  // 2, 1, 0.5, 0.25 -> codes 0..3
         } else {
            code = 3;  // every ordered quartet: weight 1/4
         }
         out.push_back(OP_SET << 30 | (uint32_t)code);
         out.push_back(pack16(i, j));
         out.push_back(pack16(k, l));
         cnt[0]++;
         const Basis *B4[4] = {&S.kinds[S.atoms[i].kind], &S.kinds[S.atoms[j].kind], &S.kinds[S.atoms[k].kind],
                               &S.kinds[S.atoms[l].kind]};
         int64_t nq = 0;
         for (auto &a : B4[0]->sh)
            for (auto &b : B4[1]->sh)
               for (auto &c : B4[2]->sh)
                  for (auto &dd : B4[3]->sh) nq += a.nl * b.nl * c.nl * dd.nl;
         cnt[3] += nq;
      }
   }
}

// first function of (shell s, contraction n) inside its atom
int fun_in_atom(const Basis &B, int s, int n)
{
   int f = 0;
   for (int t = 0; t < s; t++) f += B.sh[t].nl * (2 * B.sh[t].l + 1);
   return f + n * (2 * B.sh[s].l + 1);
}

void block_of(const GdrvSystem &S, int x, int fx, int y, int fy, unsigned *off, int *ld, bool *T)
{
   const int na = (int)S.atoms.size();
   if (x <= y) {
      const int nx = S.kinds[S.atoms[x].kind].nfun;
      *off = S.blk[(size_t)x * na + y] + (unsigned)fx + (unsigned)fy * (unsigned)nx;
      *ld = nx;
      *T = false;
   } else {
      const int ny = S.kinds[S.atoms[y].kind].nfun;
      *off = S.blk[(size_t)y * na + x] + (unsigned)fy + (unsigned)fx * (unsigned)ny;
      *ld = ny;
      *T = true;
   }
}

// ks_mix: set quartet number `nset` sends its calls in the natural order
// (a set-level layout), or with each class's calls reversed (offsets still
// affine in the contraction indices, but not the set-level layout), or with
// the first two calls of each class swapped (not affine: explicit records).
// The library pairs a call with the contraction of its position, so every
// order is a valid, different stream.
void emit_set(const GdrvSystem &S, const GdrvAbi *abi, void *h, int i, int j, int k, int l, int code,
              int64_t nset = 0)
{
   static const double fac[5] = {2.0, 1.0, 0.5, 0.25, 0.0};
   const int mix = S.p.ks_mix ? (int)(nset % 3) : 0;
   const Basis *B[4] = {&S.kinds[S.atoms[i].kind], &S.kinds[S.atoms[j].kind], &S.kinds[S.atoms[k].kind],
                        &S.kinds[S.atoms[l].kind]};
   for (size_t sa = 0; sa < B[0]->sh.size(); sa++)
      for (size_t sb = 0; sb < B[1]->sh.size(); sb++)
         for (size_t sc = 0; sc < B[2]->sh.size(); sc++)
            for (size_t sd = 0; sd < B[3]->sh.size(); sd++) {
               const Shell &A = B[0]->sh[sa], &Bb = B[1]->sh[sb], &C = B[2]->sh[sc], &D = B[3]->sh[sd];
               abi->add_qrt(h, A.l, Bb.l, C.l, D.l, A.nl, Bb.nl, C.nl, D.nl);
               const int nq = A.nl * Bb.nl * C.nl * D.nl;
               for (int n = 0; n < nq; n++) {
                  int m = n;
                  if (mix == 1) m = nq - 1 - n;
                  else if (mix == 2 && nq >= 3 && n < 2) m = 1 - n;
                  const int a = m / (Bb.nl * C.nl * D.nl), b = (m / (C.nl * D.nl)) % Bb.nl, c = (m / D.nl) % C.nl,
                            d = m % D.nl;
                  const int fa = fun_in_atom(*B[0], (int)sa, a), fb = fun_in_atom(*B[1], (int)sb, b);
                  const int fc = fun_in_atom(*B[2], (int)sc, c), fd = fun_in_atom(*B[3], (int)sd, d);
                  unsigned oac, oad, obc, obd;
                  int lac, lad, lbc, lbd;
                  bool Tac, Tad, Tbc, Tbd;
                  block_of(S, i, fa, k, fc, &oac, &lac, &Tac);
                  block_of(S, i, fa, l, fd, &oad, &lad, &Tad);
                  block_of(S, j, fb, k, fc, &obc, &lbc, &Tbc);
                  block_of(S, j, fb, l, fd, &obd, &lbd, &Tbd);
                  abi->add_qrtt(h, fac[code], A.l, Bb.l, C.l, D.l, a, b, c, d, lac, lad, lbc, lbd, oac, oad, obc, obd,
                                Tac, Tad, Tbc, Tbd);
               }
            }
   abi->add_set(h);
}

// the same set quartet in one call: the offsets of function 0 of every pair
// block; the library derives those of each (shell, contraction)
void emit_set_blocks(const GdrvSystem &S, const GdrvAbi *abi, void *h, int i, int j, int k, int l, int code)
{
   static const double fac[5] = {2.0, 1.0, 0.5, 0.25, 0.0};
   unsigned off[4];
   int ld[4];
   bool T[4];
   block_of(S, i, 0, k, 0, &off[0], &ld[0], &T[0]);
   block_of(S, i, 0, l, 0, &off[1], &ld[1], &T[1]);
   block_of(S, j, 0, k, 0, &off[2], &ld[2], &T[2]);
   block_of(S, j, 0, l, 0, &off[3], &ld[3], &T[3]);
   abi->add_set_blocks(h, fac[code], ld[0], ld[1], ld[2], ld[3], off[0], off[1], off[2], off[3], T[0], T[1], T[2],
                       T[3]);
   abi->add_set(h);
}

}  // namespace

extern "C" {

GdrvSystem *gdrv_system_new(const GdrvParams *p)
{
   GdrvSystem *S = new GdrvSystem();
   S->p = *p;
   build(*S);
   return S;
}

void gdrv_system_free(GdrvSystem *S) { delete S; }

// out: natom, nbf, P_size, Ng, nkinds, np[O], np[H], np[TM]
void gdrv_system_info(const GdrvSystem *S, int64_t out[8])
{
   out[0] = (int64_t)S->atoms.size();
   out[1] = S->nbf;
   out[2] = (int64_t)S->P_size;
   out[3] = (int64_t)S->neighs.size() / 3;
   out[4] = (int64_t)S->kinds.size();
   out[5] = S->kinds[0].np;
   out[6] = S->kinds[1].np;
   out[7] = S->kinds[2].np;
}

void gdrv_system_P(const GdrvSystem *S, double *out) { std::memcpy(out, S->P.data(), sizeof(double) * S->P_size); }
void gdrv_system_cell(const GdrvSystem *S, double *h, double *hi)
{
   std::memcpy(h, S->cell18, 9 * sizeof(double));
   std::memcpy(hi, S->cell18 + 9, 9 * sizeof(double));
}
void gdrv_system_neighs(const GdrvSystem *S, double *out) { std::memcpy(out, S->neighs.data(), sizeof(double) * S->neighs.size()); }
double gdrv_system_rcut(const GdrvSystem *S) { return S->R_cut; }
// atom positions (3 x natom), kinds, first function of each atom
void gdrv_system_atoms(const GdrvSystem *S, double *R, int *kind, int *fun0)
{
   for (size_t a = 0; a < S->atoms.size(); a++) {
      for (int k = 0; k < 3; k++) R[3 * a + k] = S->atoms[a].R[k];
      kind[a] = S->atoms[a].kind;
      fun0[a] = S->fun0[a];
   }
}
// basis of a kind: np, exponents (16), nshell, then per shell (l, nl) and
// raw coefficients concatenated
int gdrv_kind_basis(const GdrvSystem *S, int kind, double *Z, int *lnl, double *K)
{
   const Basis &B = S->kinds[kind];
   std::memcpy(Z, B.Z, sizeof(double) * B.np);
   size_t ko = 0;
   for (size_t s = 0; s < B.sh.size(); s++) {
      lnl[2 * s] = B.sh[s].l;
      lnl[2 * s + 1] = B.sh[s].nl;
      std::memcpy(K + ko, B.sh[s].K.data(), sizeof(double) * B.sh[s].K.size());
      ko += B.sh[s].K.size();
   }
   return (int)B.sh.size();
}

// set_Atom / set_Atom_L for every atom (set index = atom index)
void gdrv_emit_basis(const GdrvSystem *S, const GdrvAbi *abi, void *h)
{
   for (size_t a = 0; a < S->atoms.size(); a++) {
      const Basis &B = S->kinds[S->atoms[a].kind];
      double R[3] = {S->atoms[a].R[0], S->atoms[a].R[1], S->atoms[a].R[2]};
      double Z[16];
      std::memcpy(Z, B.Z, sizeof Z);
      abi->set_Atom(h, (int)a, R, Z, B.np);
      for (const Shell &s : B.sh) {
         std::vector<double> K = s.K;
         abi->set_Atom_L(h, (int)a, s.l, s.nl, K.data());
      }
   }
}

// the screened call stream of the system; sample > 1 keeps only the set
// quartets of the hash sample of 1 in `sample` (in_sample)
GdrvStream *gdrv_stream_record(const GdrvSystem *S, int nthreads, int64_t sample)
{
   const int na = (int)S->atoms.size();
   std::vector<std::pair<int, int>> plist;
   for (int i = 0; i < na; i++)
      for (int j = (S->p.symmetric ? i : 0); j < na; j++) plist.push_back({i, j});
   // pair lists; the smallest primitive pair that can still pass is bounded
   // by the largest partner and the largest density block
   double qall = 0, pall = 0;
   for (auto &k : S->kinds)
      for (int a = 0; a < k.np; a++) {
         const double z = 2 * k.Z[a];
         qall = std::max(qall, k.cmax_p[a] * k.cmax_p[a] * std::sqrt(2.0 * std::pow(M_PI, 2.5) / (z * z * std::sqrt(2.0 * z))));
      }
   for (double v : S->Pmax) pall = std::max(pall, v);
   const double qmin = S->p.eps / std::max(qall * pall, 1e-300);
   std::vector<std::vector<PairImage>> pairs(plist.size());
   if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 8)
   for (long t = 0; t < (long)plist.size(); t++) pairs[t] = pair_images(*S, plist[t].first, plist[t].second, qmin);
   std::vector<std::vector<uint32_t>> chunks(plist.size());
   std::vector<std::array<int64_t, 4>> counts(plist.size(), {0, 0, 0, 0});
#pragma omp parallel for schedule(dynamic, 1)
   for (long b = 0; b < (long)plist.size(); b++) {
      int64_t c[4] = {0, 0, 0, 0};
      record_bra(*S, pairs, plist, (size_t)b, chunks[b], c, sample);
      for (int q = 0; q < 4; q++) counts[b][q] = c[q];
   }
   GdrvStream *st = new GdrvStream();
   size_t tot = 0;
   for (auto &c : chunks) tot += c.size();
   st->w.reserve(tot);
   for (size_t b = 0; b < chunks.size(); b++) {
      // set start positions: scan for SET ops
      st->w.insert(st->w.end(), chunks[b].begin(), chunks[b].end());
      st->n_sets += counts[b][0];
      st->n_rows += counts[b][1];
      st->n_prm += counts[b][2];
      st->n_qrtt += counts[b][3];
      std::vector<uint32_t>().swap(chunks[b]);
   }
   // index of the first word of every set
   int64_t start = 0;
   for (size_t i = 0; i < st->w.size();) {
      const uint32_t op = st->w[i] >> 30;
      if (op == OP_PRM) {
         i++;
      } else if (op == OP_SHELL) {
         i += 3;
      } else {
         st->set_word.push_back(start);
         i += 3;
         start = (int64_t)i;
      }
   }
   return st;
}

// out: words, sets, rows (add_shell calls), primitive quartets, add_qrtt calls
void gdrv_stream_info(const GdrvStream *st, int64_t out[5])
{
   out[0] = (int64_t)st->w.size();
   out[1] = st->n_sets;
   out[2] = st->n_rows;
   out[3] = st->n_prm;
   out[4] = st->n_qrtt;
}

void gdrv_stream_free(GdrvStream *st) { delete st; }

// replay sets [set_begin, set_end) with stride into the ABI; blocks != 0:
// one add_set_blocks call per set quartet in place of its add_qrt / add_qrtt
// calls
void gdrv_stream_replay(const GdrvSystem *S, const GdrvStream *st, const GdrvAbi *abi, void *h, int64_t set_begin,
                        int64_t set_end, int64_t stride, int blocks)
{
   if (stride < 1) stride = 1;
   if (blocks && !abi->add_set_blocks) blocks = 0;
   set_end = std::min<int64_t>(set_end, (int64_t)st->set_word.size());
   for (int64_t s = set_begin; s < set_end; s += stride) {
      size_t i = (size_t)st->set_word[s];
      for (;;) {
         const uint32_t w = st->w[i];
         const uint32_t op = w >> 30;
         if (op == OP_PRM) {
            abi->add_prm(h, w & 15, (w >> 4) & 15, (w >> 8) & 15, (w >> 12) & 15);
            i++;
         } else if (op == OP_SHELL) {
            const uint32_t ij = st->w[i + 1], kl = st->w[i + 2];
            abi->add_shell(h, ij & 0xffff, ij >> 16, kl & 0xffff, kl >> 16, w & 255, (w >> 8) & 255);
            i += 3;
         } else {
            const uint32_t ij = st->w[i + 1], kl = st->w[i + 2];
            if (blocks)
               emit_set_blocks(*S, abi, h, ij & 0xffff, ij >> 16, kl & 0xffff, kl >> 16, (int)(w & 7));
            else
               emit_set(*S, abi, h, ij & 0xffff, ij >> 16, kl & 0xffff, kl >> 16, (int)(w & 7), s);
            break;
         }
      }
   }
}

// the sets (i, j, k, l) of every recorded set quartet, 4 ints each
void gdrv_stream_sets(const GdrvStream *st, int32_t *out)
{
   for (size_t s = 0; s < st->set_word.size(); s++) {
      size_t i = (size_t)st->set_word[s];
      while ((st->w[i] >> 30) != OP_SET) i += (st->w[i] >> 30) == OP_PRM ? 1 : 3;
      const uint32_t ij = st->w[i + 1], kl = st->w[i + 2];
      out[4 * s + 0] = (int32_t)(ij & 0xffff);
      out[4 * s + 1] = (int32_t)(ij >> 16);
      out[4 * s + 2] = (int32_t)(kl & 0xffff);
      out[4 * s + 3] = (int32_t)(kl >> 16);
   }
}

// replay sets [set_begin, set_end) with stride, asking for the integrals of
// every class instead of K updates: class (sa,sb,sc,sd) of each set quartet
// in turn (shells in basis order) goes to *dest, which advances by its
// nla nlb nlc nld Ns integrals
void gdrv_stream_replay_tra(const GdrvSystem *S, const GdrvStream *st, const GdrvAbi *abi, void *h, int64_t set_begin,
                            int64_t set_end, int64_t stride, uint64_t *dest)
{
   if (stride < 1) stride = 1;
   set_end = std::min<int64_t>(set_end, (int64_t)st->set_word.size());
   for (int64_t s = set_begin; s < set_end; s += stride) {
      size_t i = (size_t)st->set_word[s];
      for (;;) {
         const uint32_t w = st->w[i];
         const uint32_t op = w >> 30;
         if (op == OP_PRM) {
            abi->add_prm(h, w & 15, (w >> 4) & 15, (w >> 8) & 15, (w >> 12) & 15);
            i++;
         } else if (op == OP_SHELL) {
            const uint32_t ij = st->w[i + 1], kl = st->w[i + 2];
            abi->add_shell(h, ij & 0xffff, ij >> 16, kl & 0xffff, kl >> 16, w & 255, (w >> 8) & 255);
            i += 3;
         } else {
            const uint32_t ij = st->w[i + 1], kl = st->w[i + 2];
            const int at[4] = {(int)(ij & 0xffff), (int)(ij >> 16), (int)(kl & 0xffff), (int)(kl >> 16)};
            const Basis *B[4];
            for (int t = 0; t < 4; t++) B[t] = &S->kinds[S->atoms[at[t]].kind];
            for (auto &a : B[0]->sh)
               for (auto &b : B[1]->sh)
                  for (auto &c : B[2]->sh)
                     for (auto &d : B[3]->sh) {
                        abi->add_tra(h, a.l, b.l, c.l, d.l, *dest);
                        *dest += (uint64_t)(a.nl * b.nl * c.nl * d.nl) *
                                 (uint64_t)((2 * a.l + 1) * (2 * b.l + 1) * (2 * c.l + 1) * (2 * d.l + 1));
                     }
            abi->add_set(h);
            break;
         }
      }
   }
}

// the rest of the ABI a CP2K-like OpenMP team calls (reference
// src/interface_libgint.F90, src/extern_functions.cpp)
typedef struct {
   void *(*create_handle)(void);
   void (*init)(void *);
   void (*set_Potential_Truncated)(void *, double, double *, int, int);
   void (*set_hf_fac)(void *, double);
   void (*set_max_mem)(void *, int);
   void (*set_cell)(void *, bool, double *, double *);
   void (*set_neighs)(void *, double *, int);
   void (*set_P)(void *, double *, int);
   void (*get_K)(void *, double *);
   void (*destroy_handle)(void *);
   void (*plan_begin)(void *);   // optional (null: the stream is executed directly)
   void (*plan_end)(void *);
   void (*plan_execute)(void *);
} GdrvTeamAbi;

// One K-build by an OpenMP team of nthreads, as CP2K drives the library: a
// handle per thread (reference src/extern_functions.cpp:8), libgint_init,
// set_P and get_K called by every thread of the team (they are collective:
// the device P and K are shared, src/libGint.cpp:37-60,441-447,613-635),
// and thread t sending set quartets t, t + nthreads, ... of the stream.
// With plan != 0 each thread records its slice as a resident plan and
// executes it.  K: the team's result (written once); secs[0] the wall time
// of the team's ingestion (slowest thread), secs[1] of the whole K-build.
int gdrv_team_kbuild(const GdrvSystem *S, const GdrvStream *st, const GdrvAbi *abi, const GdrvTeamAbi *t,
                     int nthreads, int blocks, int plan, double *C0, int ld_C0, int c0_size, int max_mem, double *K,
                     double *secs)
{
   if (nthreads < 1) return -1;
   double ing = 0.0;
   const double t0 = omp_get_wtime();
   std::vector<double> P = S->P;
   double h9[9], hi9[9];
   std::memcpy(h9, S->cell18, sizeof h9);
   std::memcpy(hi9, S->cell18 + 9, sizeof hi9);
   std::vector<double> g = S->neighs;
   const int nset = (int)st->set_word.size();
#pragma omp parallel num_threads(nthreads) reduction(max : ing)
   {
      void *h = t->create_handle();
      t->init(h);
      if (C0) t->set_Potential_Truncated(h, S->R_cut, C0, ld_C0, c0_size);
      t->set_hf_fac(h, 1.0);
      t->set_max_mem(h, max_mem);
      t->set_cell(h, S->p.image_shells > 0, h9, hi9);
      t->set_neighs(h, g.data(), (int)g.size() / 3);
      gdrv_emit_basis(S, abi, h);
      t->set_P(h, P.data(), (int)P.size());
      const int me = omp_get_thread_num(), nt = omp_get_num_threads();
      const double ti = omp_get_wtime();
      if (plan && t->plan_begin) t->plan_begin(h);
      gdrv_stream_replay(S, st, abi, h, me, nset, nt, blocks);
      if (plan && t->plan_begin) t->plan_end(h);
      ing = omp_get_wtime() - ti;
      if (plan && t->plan_begin) t->plan_execute(h);
      t->get_K(h, K);
#pragma omp barrier
      t->destroy_handle(h);
   }
   secs[0] = ing;
   secs[1] = omp_get_wtime() - t0;
   return 0;
}

// symmetric part of K accumulated by the library, in the block layout:
// diagonal blocks (B + B^T)/2, off-diagonal blocks hold K_ab + K_ba^T -> /2
void gdrv_symmetrize(const GdrvSystem *S, double *K)
{
   const int na = (int)S->atoms.size();
   for (int a = 0; a < na; a++)
      for (int b = a; b < na; b++) {
         const int nx = S->kinds[S->atoms[a].kind].nfun, ny = S->kinds[S->atoms[b].kind].nfun;
         double *B = K + S->blk[(size_t)a * na + b];
         if (a == b) {
            for (int r = 0; r < nx; r++)
               for (int c = r; c < nx; c++) {
                  const double v = 0.5 * (B[r + (size_t)c * nx] + B[c + (size_t)r * nx]);
                  B[r + (size_t)c * nx] = B[c + (size_t)r * nx] = v;
               }
         } else {
            for (size_t t = 0; t < (size_t)nx * ny; t++) B[t] *= 0.5;
         }
      }
}

// dense nbf x nbf (row-major) from the block layout of a symmetric matrix
void gdrv_dense(const GdrvSystem *S, const double *M, double *D)
{
   const int na = (int)S->atoms.size(), n = S->nbf;
   for (int a = 0; a < na; a++)
      for (int b = a; b < na; b++) {
         const int nx = S->kinds[S->atoms[a].kind].nfun, ny = S->kinds[S->atoms[b].kind].nfun;
         const double *B = M + S->blk[(size_t)a * na + b];
         for (int r = 0; r < nx; r++)
            for (int c = 0; c < ny; c++) {
               const double v = B[r + (size_t)c * nx];
               D[(size_t)(S->fun0[a] + r) * n + S->fun0[b] + c] = v;
               D[(size_t)(S->fun0[b] + c) * n + S->fun0[a] + r] = v;
            }
      }
}

}  // extern "C"
