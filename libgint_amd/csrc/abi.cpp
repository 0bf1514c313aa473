// extern "C" entry points (include/libGint.h): thin pass-through to the
// engine, plus the contraction-coefficient normalisation that the reference
// performs on the C side (src/extern_functions.cpp:92-136).
#include <cmath>
#include <vector>

#include "../../include/libGint.h"
#include "engine.hpp"
#include "gint_math.h"

using gint::Engine;

static Engine *E(void *h)
{
   if (!h) gint::fatal("null libgint handle", 0);
   return static_cast<Engine *>(h);
}

extern "C" {

void *libgint_create_handle(void) { return new Engine(); }
void libgint_init(void *h) { E(h)->init(); }
void libgint_set_Potential_Truncated(void *h, double R_cut, double *C0, int ld_C0, int C0_size)
{
   E(h)->set_potential_truncated(R_cut, C0, ld_C0, C0_size);
}
void libgint_set_hf_fac(void *h, double fac) { E(h)->set_hf_fac(fac); }
void libgint_set_max_mem(void *h, int max_mem) { E(h)->set_max_mem(max_mem); }
void libgint_set_P(void *h, double *P, int P_size) { E(h)->set_P(P, P_size); }
void libgint_set_P_polarized(void *h, double *Pa, double *Pb, int P_size) { E(h)->set_P(Pa, Pb, P_size); }
void libgint_set_K(void *h, double *K, int K_size) { E(h)->set_K(K, K_size); }
void libgint_set_K_polarized(void *h, double *Ka, double *Kb, int K_size) { E(h)->set_K(Ka, Kb, K_size); }
void libgint_get_K(void *h, double *K) { E(h)->get_K(K); }
void libgint_get_K_polarized(void *h, double *Ka, double *Kb) { E(h)->get_K(Ka, Kb); }
// `periodic` is not used: the minimum-image wrap is always applied, as in
// reference src/compute_Fm.cpp:79,88,97 (the Fortran binding passes it by
// reference, so its value is not reliable anyway)
void libgint_set_cell(void *h, bool periodic, double *cell_h, double *cell_i)
{
   (void)periodic;
   E(h)->set_cell(cell_h, cell_i);
}
void libgint_set_neighs(void *h, double *neighs, int nneighs) { E(h)->set_neighs(neighs, nneighs); }
void libgint_set_Atom(void *h, int i, double *R, double *Z, int np) { E(h)->set_atom(i, R, Z, np); }

// (pi^{3/2} / 2^l) sum_ij K_i K_j / (z_i + z_j)^{3/2 + l}, square root;
// reference src/extern_functions.cpp:95-110
double compute_norm_psi(double *K, double *z, int np, int l)
{
   double ans = 0.0;
   for (int i = 0; i < np; i++)
      for (int j = 0; j < np; j++) ans += K[i] * K[j] / std::pow(z[i] + z[j], 1.5 + l);
   ans *= std::pow(M_PI, 1.5) / std::pow(2.0, l);
   return std::sqrt(ans);
}

void libgint_set_Atom_L(void *h, int i, int l, int nl, double *K)
{
   Engine *e = E(h);
   const int np = e->np_of(i);
   if (np <= 0) gint::fatal("set_Atom_L before set_Atom", i);
   std::vector<double> z(e->exponents_of(i), e->exponents_of(i) + np);
   std::vector<double> Kn((size_t)nl * np);
   for (int inl = 0; inl < nl; inl++) {
      const double norm = compute_norm_psi(&K[inl * np], z.data(), np, l);
      for (int ip = 0; ip < np; ip++) Kn[(size_t)inl * np + ip] = K[inl * np + ip] / norm;
   }
   e->set_atom_L(i, l, nl, Kn.data());
}

// reference src/extern_functions.cpp:139-147, with the shells lmin..Lmax
// read as nl[l - lmin] contractions each (the reference indexes nl by the
// set index there and stops before Lmax; its caller no longer uses it)
void libgint_set_AtomInfo(void *h, int i, double *R, double *Z, int np, int lmin, int Lmax, int *nl, double *K)
{
   libgint_set_Atom(h, i, R, Z, np);
   int k_offset = 0;
   for (int l = lmin; l <= Lmax; l++) {
      libgint_set_Atom_L(h, i, l, nl[l - lmin], &K[k_offset]);
      k_offset += nl[l - lmin] * np;
   }
}

void libgint_add_prm(void *h, int ipa, int ipb, int ipc, int ipd) { E(h)->add_prm(ipa, ipb, ipc, ipd); }
void libgint_add_shell(void *h, int i, int j, int k, int l, int n1, int n2) { E(h)->add_shell(i, j, k, l, n1, n2); }
void libgint_add_cell(void *h) { (void)h; }
void libgint_add_qrt(void *h, int la, int lb, int lc, int ld, int nla, int nlb, int nlc, int nld)
{
   E(h)->add_qrt(la, lb, lc, ld, nla, nlb, nlc, nld);
}
void libgint_add_qrtt(void *h, double symm_fac, int la, int lb, int lc, int ld, int inla, int inlb, int inlc,
                      int inld, int ld_ac, int ld_ad, int ld_bc, int ld_bd, unsigned int off_ac, unsigned int off_ad,
                      unsigned int off_bc, unsigned int off_bd, bool Tac, bool Tad, bool Tbc, bool Tbd)
{
   E(h)->add_qrtt(symm_fac, la, lb, lc, ld, inla, inlb, inlc, inld, ld_ac, ld_ad, ld_bc, ld_bd, off_ac, off_ad,
                  off_bc, off_bd, Tac, Tad, Tbc, Tbd);
}
void libgint_add_set(void *h) { E(h)->add_set(); }
void libgint_amd_add_set_blocks(void *h, double symm_fac, int ld_ac, int ld_ad, int ld_bc, int ld_bd, unsigned off_ac,
                                unsigned off_ad, unsigned off_bc, unsigned off_bd, bool Tac, bool Tad, bool Tbc, bool Tbd)
{
   const int ld[4] = {ld_ac, ld_ad, ld_bc, ld_bd};
   const unsigned off[4] = {off_ac, off_ad, off_bc, off_bd};
   const bool T[4] = {Tac, Tad, Tbc, Tbd};
   E(h)->add_set_blocks(symm_fac, ld, off, T);
}
void libgint_dispatch(void *h) { E(h)->dispatch_all(); }

void libgint_amd_set_output(void *h, double *out, uint64_t n) { E(h)->set_output(out, n); }
void libgint_amd_add_tra(void *h, int la, int lb, int lc, int ld, uint64_t dest) { E(h)->add_tra(la, lb, lc, ld, dest); }
void libgint_amd_get_output(void *h) { E(h)->get_output(); }
void libgint_amd_destroy_handle(void *h) { delete static_cast<Engine *>(h); }
void libgint_amd_synchronize(void *h) { E(h)->synchronize(); }
void libgint_amd_enable_stats(void *h, int on) { E(h)->enable_stats(on != 0); }
void libgint_amd_stats(void *h, uint64_t out[8]) { E(h)->stats(out); }
void libgint_amd_enable_timing(void *h, int on) { E(h)->enable_timing(on != 0); }
void libgint_amd_timing(void *h, double out[12])
{
   const gint::Timing t = E(h)->timing();
   for (int i = 0; i < 4; i++) {
      out[i] = t.ms[i];
      out[4 + i] = (double)t.launches[i];
      out[8 + i] = t.flops[i];
   }
}
void libgint_amd_timing_eco(void *h, double out[3])
{
   const gint::Timing t = E(h)->timing();
   out[0] = t.ms[4];
   out[1] = (double)t.launches[4];
   out[2] = t.flops[4];
}
void libgint_amd_reset_counters(void *h) { E(h)->reset_counters(); }
void libgint_amd_timing_detail(void *h, int cat, double out[256]) { E(h)->timing_detail(cat, out); }
void libgint_amd_set_debug(void *h, int flags) { E(h)->set_debug(flags); }
void libgint_amd_set_graph(void *h, int on) { E(h)->set_graph(on != 0); }
void libgint_amd_set_autotune(void *h, int on) { E(h)->set_autotune(on != 0); }
void libgint_amd_set_mixed_precision(void *h, double tau) { E(h)->set_mixed_precision(tau); }
void libgint_amd_mixed_info(void *h, double out[4]) { E(h)->mixed_info(out); }
void libgint_amd_tune_info(void *h, double out[5]) { E(h)->tune_info(out); }
void libgint_amd_tune_detail(void *h, double ms[256], int pick[256]) { E(h)->tune_detail(ms, pick); }
void libgint_amd_graph_info(void *h, uint64_t out[4]) { E(h)->graph_info(out); }
void libgint_amd_plan_begin(void *h) { E(h)->plan_begin(); }
void libgint_amd_plan_end(void *h) { E(h)->plan_end(); }
void libgint_amd_plan_execute(void *h) { E(h)->plan_execute(); }
int libgint_amd_plan_execute_noalloc(void *h, const double *P, const double *Pb, double *K, double *Kb, uint64_t n)
{
   return E(h)->plan_execute_noalloc(P, Pb, K, Kb, n);
}
void libgint_amd_plan_clear(void *h) { E(h)->plan_clear(); }
void libgint_amd_plan_info(void *h, uint64_t out[6]) { E(h)->plan_info(out); }
void libgint_amd_queue_info(void *h, uint64_t out[8]) { E(h)->queue_info(out); }
int libgint_amd_queue_types(void *h, uint64_t *out, int max_types) { return E(h)->queue_types(out, max_types); }
int libgint_amd_mfma_selftest(const double *A, const double *B, double *D, int prec)
{
   return gint::mfma_selftest(A, B, D, prec);
}
int libgint_amd_math_selftest(const double *x, const double *y, int n, double *out)
{
   return gint::device_math_selftest(x, y, n, out);
}
void libgint_amd_math_host(const double *x, const double *y, int n, double *out)
{
   for (int i = 0; i < n; i++) {
      const double a = x[i], c = y[i];
      out[4 * i + 0] = gint_exp(a);
      out[4 * i + 1] = std::sqrt(std::fabs(a));
      out[4 * i + 2] = 1. / a;
      out[4 * i + 3] = a * c / (a + c);
   }
}

}  // extern "C"
