/*
 * CPU oracle for the libGint K-build -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of the reference engine's algorithm, used by
 * the tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg
 * as the checker.  Nothing in the product library links or calls it.
 *
 * The entry points mirror the libgint_* C-ABI (same arguments, same call
 * protocol) so that one call stream can be replayed into both.  Each set
 * quartet is integrated when libgint-style add_set() arrives, so the oracle
 * needs the density (set_P) before the first add_set, like the reference.
 */
#ifndef LIBGINT_ORACLE_H
#define LIBGINT_ORACLE_H

#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

void *oracle_create_handle(void);
void oracle_destroy_handle(void *h);
void oracle_init(void *h);
void oracle_set_Potential_Truncated(void *h, double R_cut, double *C0, int ld_C0, int C0_size);
void oracle_set_hf_fac(void *h, double fac);
void oracle_set_max_mem(void *h, int max_mem_MiB);
void oracle_set_P(void *h, double *P, int P_size);
void oracle_set_P_polarized(void *h, double *Pa, double *Pb, int P_size);
void oracle_set_K(void *h, double *K, int K_size);
void oracle_set_K_polarized(void *h, double *Ka, double *Kb, int K_size);
void oracle_get_K(void *h, double *K);
void oracle_get_K_polarized(void *h, double *Ka, double *Kb);
void oracle_set_cell(void *h, bool periodic, double *cell_h, double *cell_i);
void oracle_set_neighs(void *h, double *neighs, int nneighs);
void oracle_set_Atom(void *h, int i, double *R, double *Z, int np);
void oracle_set_Atom_L(void *h, int i, int l, int nl, double *K);
void oracle_add_prm(void *h, int ipa, int ipb, int ipc, int ipd);
void oracle_add_shell(void *h, int i, int j, int k, int l, int n1, int n2);
void oracle_add_cell(void *h);
void oracle_add_qrt(void *h, int la, int lb, int lc, int ld, int nla, int nlb, int nlc, int nld);
void oracle_add_qrtt(void *h, double symm_fac, int la, int lb, int lc, int ld,
                     int inla, int inlb, int inlc, int inld, int ld_ac, int ld_ad, int ld_bc, int ld_bd,
                     unsigned off_ac, unsigned off_ad, unsigned off_bc, unsigned off_bd,
                     bool Tac, bool Tad, bool Tbc, bool Tbd);
void oracle_add_set(void *h);
/* the add_qrt / add_qrtt calls of the open set quartet for K blocks laid out
   per pair of sets (the product's libgint_amd_add_set_blocks) */
void oracle_add_set_blocks(void *h, double symm_fac, int ld_ac, int ld_ad, int ld_bc, int ld_bd,
                           unsigned off_ac, unsigned off_ad, unsigned off_bc, unsigned off_bd,
                           bool Tac, bool Tad, bool Tbc, bool Tbd);
void oracle_dispatch(void *h);

/* integral output in the reference's TRA layout (the product's
   libgint_amd_set_output / add_tra / get_output); set_output zeroes out */
void oracle_set_output(void *h, double *out, uint64_t n);
void oracle_add_tra(void *h, int la, int lb, int lc, int ld, uint64_t dest);
void oracle_get_output(void *h);

/* normalisation of one contraction, reference src/extern_functions.cpp:95-110 */
double oracle_norm_psi(const double *K, const double *z, int np, int l);

/* statistics of the slots integrated so far:
   out[0] slots evaluated (primitive quartet x cell), out[1] slots whose
   Fm[0] > 1e-12 ran the recurrences, out[2] primitive quartets,
   out[3] contracted quartets (add_qrtt entries), out[4] order-independent
   64-bit checksum of the surviving slots (gint_slot_key). */
void oracle_stats(void *h, uint64_t out[5]);

/* pieces of the chain, exposed for unit tests */
double oracle_exp(double x);
void oracle_boys(int nmax, double T, double *f, double fac);
/* G_m(R,T) m=0..nder from the C0 table loaded by set_Potential_Truncated;
   returns the branch taken (0 zero, 1 Boys, 2 patch) */
int oracle_tcg(void *h, double R, double T, int nder, double *G);
void oracle_patch_grid(int *grid_lo, int *grid_hi);
void oracle_patch_bw(double *bw /* 207*4 */);
void oracle_c2s(double *out /* 110 */);
void oracle_boys_table(double *ftable /* 121*28 */);

/* contracted cartesian integrals [nlabcd][Nc] (layout [a][b][c][d]) of
   one row; prm packs ipa | ipb<<8 | ipc<<16 | ipd<<24 */
void oracle_row_cart(void *h, int i, int j, int k, int l, int n1, int n2, int sa, int sb, int sc, int sd,
                     int nprm, const uint32_t *prm, double *out);

/* keys for the order-independent survivor checksum */
uint64_t oracle_slot_key(int i, int j, int k, int l, int n1, int n2,
                         int la, int lb, int lc, int ld,
                         int ipa, int ipb, int ipc, int ipd, int n3);

#ifdef __cplusplus
}
#endif
#endif
