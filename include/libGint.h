/*
 * libGint C-ABI for the MI355X (gfx950) K-build.
 *
 * The first group of entry points is the reference library's exported
 * interface, one for one: same names, argument order and meaning, bound by
 * CP2K through the reference's Fortran module (reference
 * src/interface_libgint.F90:19-241).  Each declaration cites the reference
 * function it replaces (src/extern_functions.cpp).  The second group is
 * this library's own and is marked as such.
 *
 * Protocol (reference src/interface_libgint.F90:404-462): per set quartet,
 * { {libgint_add_prm}* libgint_add_shell }* over the (n1,n2) images, then
 * per angular-momentum combination libgint_add_qrt followed by nla*nlb*nlc*nld
 * libgint_add_qrtt calls in (inla,inlb,inlc,inld) row-major order, then
 * libgint_add_set.  libgint_init, libgint_set_P*, libgint_set_K* and
 * libgint_get_K* are collective over the caller's OpenMP team.
 *
 * Errors: the reference returns void and exits on device errors; this
 * library keeps the void signatures, prints a message and aborts on device
 * errors and on arguments beyond its limits (l <= 3, <= 16 primitives per
 * set, <= 4 contractions per shell, <= 256 lattice vectors).
 */
#ifndef LIBGINT_AMD_LIBGINT_H
#define LIBGINT_AMD_LIBGINT_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference interface ---------------------------------------------- */

/* src/extern_functions.cpp:9-15: new engine, one per OpenMP thread */
void *libgint_create_handle(void);
/* src/extern_functions.cpp:17-20: per-thread stream set-up (collective) */
void libgint_init(void *handle);
/* src/extern_functions.cpp:22-25: truncated Coulomb, C0 of 207 patches x ld_C0 */
void libgint_set_Potential_Truncated(void *handle, double R_cut, double *C0, int ld_C0, int C0_size);
/* src/extern_functions.cpp:27-32: K += fac * exchange */
void libgint_set_hf_fac(void *handle, double fac);
/* src/extern_functions.cpp:34-37: device memory budget per MPI rank, MiB */
void libgint_set_max_mem(void *handle, int max_mem);
/* src/extern_functions.cpp:40-44: density upload, also zeroes K (collective) */
void libgint_set_P(void *handle, double *P, int P_size);
/* src/extern_functions.cpp:46-50 */
void libgint_set_P_polarized(void *handle, double *Pa, double *Pb, int P_size);
/* src/extern_functions.cpp:52-55: start K from the given values (collective) */
void libgint_set_K(void *handle, double *K, int K_size);
/* src/extern_functions.cpp:57-60 */
void libgint_set_K_polarized(void *handle, double *Ka, double *Kb, int K_size);
/* src/extern_functions.cpp:62-65: finish all work, copy K to the host (collective) */
void libgint_get_K(void *handle, double *K);
/* src/extern_functions.cpp:67-70 */
void libgint_get_K_polarized(void *handle, double *Ka, double *Kb);
/* src/extern_functions.cpp:72-75: h and h^-1, 3x3 column-major, lattice vectors as columns */
void libgint_set_cell(void *handle, bool periodic, double *cell_h, double *cell_i);
/* src/extern_functions.cpp:77-80: lattice vectors (3 x nneighs, column-major) */
void libgint_set_neighs(void *handle, double *neighs, int nneighs);
/* src/extern_functions.cpp:82-88: set i at R with np exponents Z */
void libgint_set_Atom(void *handle, int i, double *R, double *Z, int np);
/* src/extern_functions.cpp:95-110: norm used to normalise one contraction */
double compute_norm_psi(double *K, double *z, int np, int l);
/* src/extern_functions.cpp:114-136: shell l of set i, nl x np coefficients */
void libgint_set_Atom_L(void *handle, int i, int l, int nl, double *K);
/* src/extern_functions.cpp:139-147: unused by the reference's caller; kept */
void libgint_set_AtomInfo(void *handle, int i, double *R, double *Z, int np, int lmin, int Lmax, int *nl, double *K);
/* src/extern_functions.cpp:149-152: one primitive quartet of the next add_shell */
void libgint_add_prm(void *handle, int ipa, int ipb, int ipc, int ipd);
/* src/extern_functions.cpp:154-157: set quartet (i,j,k,l) at images n1 (b) and n2 (d) */
void libgint_add_shell(void *handle, int i, int j, int k, int l, int n1, int n2);
/* src/extern_functions.cpp:159-162: no-op */
void libgint_add_cell(void *handle);
/* src/extern_functions.cpp:164-167: declares nla*nlb*nlc*nld quartets of class (la,lb,lc,ld) */
void libgint_add_qrt(void *handle, int la, int lb, int lc, int ld, int nla, int nlb, int nlc, int nld);
/* src/extern_functions.cpp:169-180: where one contracted quartet goes in K */
void libgint_add_qrtt(void *handle, double symm_fac, int la, int lb, int lc, int ld, int inla, int inlb, int inlc,
                      int inld, int ld_ac, int ld_ad, int ld_bc, int ld_bd, unsigned int offset_ac_L_set,
                      unsigned int offset_ad_L_set, unsigned int offset_bc_L_set, unsigned int offset_bd_L_set,
                      bool Tac, bool Tad, bool Tbc, bool Tbd);
/* src/extern_functions.cpp:182-185: closes a set quartet */
void libgint_add_set(void *handle);
/* src/extern_functions.cpp:187-190: run all pending work now */
void libgint_dispatch(void *handle);

/* ---- this library's own (no counterpart in the reference) -------------- */

/* Integral output, the use of reference src/compute_TRA.cpp:8-56 (whose
   launch the reference's K path leaves out, src/libGint.cpp:992): after
   libgint_amd_set_output(h, out, n) the handle keeps an n-double output
   array on the device, zeroed; libgint_amd_add_tra(h, la, lb, lc, ld, dest),
   called for the current set quartet in place of that class's add_qrt /
   add_qrtt calls, sends the class's spherical integrals (every contraction
   combination) to out[dest ...] in the reference's TRA order
   [ila, isa, ilb, isb, ilc, isc, ild, isd] (nla*nlb*nlc*nld*Ns doubles,
   written, not added; no symm_fac or hf_fac factor); the other classes of
   the set quartet keep their add_qrtt calls.  libgint_amd_get_output
   finishes the work and copies the array to `out`.  set_output(h, NULL, 0)
   releases it.  A class of a set quartet with no surviving primitive
   quartet writes nothing (its integrals are zero). */
void libgint_amd_set_output(void *handle, double *out, uint64_t n);
void libgint_amd_add_tra(void *handle, int la, int lb, int lc, int ld, uint64_t dest);
void libgint_amd_get_output(void *handle);

/* release a handle (the reference never frees one) */
void libgint_amd_destroy_handle(void *handle);
/* wait for the handle's device work */
void libgint_amd_synchronize(void *handle);
/* count surviving (primitive quartet, cell) slots and their checksum */
void libgint_amd_enable_stats(void *handle, int on);
/* out[0] (primitive quartet, cell, class) slots evaluated, [1] slots with
   Fm[0] > 1e-12 (per class, the reference's recurrence work unit),
   [2] primitive quartets x classes, [3] add_qrtt calls, [4] checksum of the
   surviving slots, [5] rows (add_shell calls with primitives), [6] algorithmic
   FLOPs of the fm_r kernel: the reference chain's Fm + VRR FLOPs of the
   (slot, class) pairs kept, [7] 0 */
void libgint_amd_stats(void *handle, uint64_t out[8]);
/* per-kernel device time with HIP events on the handle's stream */
void libgint_amd_enable_timing(void *handle, int on);
/* out[0..3] ms in fm_r / the wavefront-per-set-quartet kernels tq /
   the per-set-quartet quartet kernel gqs / fix-up, out[4..7] launches,
   out[9], out[10] algorithmic FLOPs of tq and gqs (out[8] and out[11] = 0:
   the fm_r FLOPs are libgint_amd_stats out[6]) */
void libgint_amd_timing(void *handle, double out[12]);
/* the matrix-core contraction of the classes with many contractions (done
   before the per-class kernel, which then reads its result): out[0] ms,
   [1] launches, [2] algorithmic FLOPs (2 n_prm ncp per contraction) */
void libgint_amd_timing_eco(void *handle, double out[3]);
void libgint_amd_reset_counters(void *handle);
/* ms per launch key of one category of libgint_amd_timing (0 fm_r, 1 tq,
   2 gqs, 3 fix-up, 4 matrix-core contraction) by the quartet type's total
   angular momentum */
void libgint_amd_timing_detail(void *handle, int cat, double out[256]);
/* bit 0: route every surviving slot through the class-dependent screening
   path (a test of that path; results are unchanged); bit 1: run every
   kernel of a chunk on the handle's stream, one after the other (per-kernel
   durations without overlap, for profiling; results are unchanged); bit 2:
   wait for the work of each quartet type and report its completion on
   stderr (progress of long K-builds; results are unchanged); bit 3: every
   class that can take it goes to the Kronecker HRR, whatever the autotuner
   chose (resident plans) or the default rule (call streams), with the matrix-core contraction where the class
   has one (else its own); bit 4: the same with the class's own contraction
   where it fits (tests of that path; results are unchanged) */
void libgint_amd_set_debug(void *handle, int flags);
/* libgint_amd_plan_execute replays the plan's launches (every chunk's
   kernels on the handle's stream and its side streams) as one HIP graph:
   captured at the first execution, replayed at the next ones while the
   buffers, P / K arrays, potential and debug flags it was captured with are
   unchanged (else captured again).  Executions with timing, statistics or
   debug bits 1-2 on, and handles of an OpenMP team of more than one thread
   (whose other threads make allocating and synchronous calls while one
   would capture), run eagerly.  on = 0 turns the graph off (default on). */
void libgint_amd_set_graph(void *handle, int on);
/* out[0] graphs on, [1] captures, [2] replays, [3] a graph is held */
void libgint_amd_graph_info(void *handle, uint64_t out[4]);
/* Autotuner (default on): at the first libgint_amd_plan_execute of a plan,
   each class of each quartet type gets the fastest of its schedules, timed
   on the type's first set quartets (into a scratch K): the templated
   one-wavefront kernel, the workgroup kernel with its own contraction, or
   the matrix-core contraction followed by the workgroup kernel.  Decisions
   are kept per (class, contractions, primitive quartets per row in powers
   of two) for the process and reused.  on = 0 keeps the built-in choice. */
void libgint_amd_set_autotune(void *handle, int on);
/* Mixed precision for resident plans (tau = 0, the default: FP64 only).
   The first execution after plan_end (or after tau is first set) runs in
   FP64 and records each set quartet's largest primitive block element;
   later executions send a set quartet whose bound -- that element times
   its largest contraction coefficients (one per centre) times its
   primitive quartets times max |P| of the density -- is below tau to the
   single-precision matrix-core contraction (each batch of 16 primitive
   quartets summed in single precision, batches, HRR, spherical transform
   and K in double precision).  The Boys function, the Hermite recursion
   and the other classes stay FP64. */
void libgint_amd_set_mixed_precision(void *handle, double tau);
/* out[0] tau, [1] calibrated, [2] set quartets routed to single precision
   by the last execution, [3] set quartets of the plan */
void libgint_amd_mixed_info(void *handle, double out[4]);
/* out[0] the plan is tuned, [1..3] classes given the templated kernel /
   the workgroup kernel with its own contraction / the matrix-core
   contraction (each with either HRR), [4] seconds spent */
void libgint_amd_tune_info(void *handle, double out[5]);
/* per class code ((la*4+lb)*4+lc)*4+ld: ms[] the chosen schedule's time on
   the timing samples (summed over quartet types), pick[] that schedule
   (0 templated, 1 workgroup, 2 matrix-core contraction + layered HRR,
   3 matrix-core contraction + Kronecker HRR, 4 workgroup kernel's own
   contraction + Kronecker HRR; last type timed) */
void libgint_amd_tune_detail(void *handle, double ms[256], int pick[256]);
/* resident plans: the calls between plan_begin and plan_end are kept on the
   device; plan_execute adds their exchange terms to K for the density of the
   last set_P (between set_P and get_K, in place of the calls) */
void libgint_amd_plan_begin(void *handle);
void libgint_amd_plan_end(void *handle);
void libgint_amd_plan_execute(void *handle);
void libgint_amd_plan_clear(void *handle);
/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): the handle's resident plan, as libgint_amd_plan_execute
   runs it, but with the density read from the caller's device array P
   (Pb: spin b, or NULL) and the exchange terms ADDED into the caller's
   device array K (Kb), n doubles each (K is not zeroed).  Every scratch
   buffer of the kernels (primitive blocks, fix-up list, class workspace,
   counters) is carved from one static device array of 176 MiB per GPU,
   held until the process ends; the call allocates and frees nothing.  It
   is asynchronous on the handle's stream (libgint_amd_synchronize waits).
   Calls on one GPU share that array, which keeps the last call's contents:
   calls that overlap in time on one GPU (other handles, threads or streams)
   race and give wrong results, so serialise them.  Returns 0, or, with a
   message on stderr and nothing launched: -1 protocol (recording, no
   libgint_init, an open set quartet), -2 missing arrays, -3 the plan's
   offsets reach beyond n, -4 one set quartet's primitive blocks exceed the
   array's 130 MiB block region (libgint_amd_plan_execute takes those),
   -5 a class workspace beyond its 32 MiB region, -6 tables or basis data
   changed since libgint_amd_plan_end, -7 the C0 table has too few orders,
   -8 the array has no device address. */
int libgint_amd_plan_execute_noalloc(void *handle, const double *P, const double *Pb, double *K, double *Kb,
                                     uint64_t n);
/* out[0] resident bytes, [1] quartet types, [2] chunks, [3] set quartets,
   [4] rows, [5] primitive quartets */
void libgint_amd_plan_info(void *handle, uint64_t out[6]);
/* work recorded and not yet on the device (host only, no GPU call):
   out[0] quartet types, [1] set quartets, [2] rows, [3] primitive quartets,
   [4] (set quartet, class) pairs with add_qrtt calls, [5] of them kept as
   explicit lists because their offsets are not affine, [6] set quartets
   whose calls reduce to one set-level record (libgint_amd_add_set_blocks),
   [7] bytes of the largest set quartet's primitive blocks */
void libgint_amd_queue_info(void *handle, uint64_t out[8]);
/* the same per quartet type with recorded work, 8 numbers each: total
   angular momentum L, primitive-block doubles ncp, classes, set quartets,
   rows, primitive quartets, contracted quartets per set quartet, shells per
   set (4-bit fields); fills at most max_types and returns the number of
   types (host only, no GPU call) */
int libgint_amd_queue_types(void *handle, uint64_t *out, int max_types);
/* Every add_qrt / add_qrtt call of the current set quartet (i,j,k,l) at
   once, for a caller whose P / K blocks are laid out per pair of sets, as
   CP2K's block matrices are: element (f, g) of the (a,c) pair block, f a
   spherical function of set i and g one of set k, is at
   off_ac + (Tac ? f * ld_ac + g : g * ld_ac + f), likewise for (a,d), (b,c)
   and (b,d); the functions of a set run shell by shell in libgint_set_Atom_L
   order, contraction by contraction, 2l+1 each.  Equivalent to add_qrt and
   all nla*nlb*nlc*nld add_qrtt calls of every class of the set quartet with
   those offsets and symm_fac; call it after the set quartet's add_shell
   calls, then libgint_add_set.  A per-call stream with this layout is
   recognised at libgint_add_set and stored the same way (32 bytes per set
   quartet). */
void libgint_amd_add_set_blocks(void *handle, double symm_fac, int ld_ac, int ld_ad, int ld_bc, int ld_bd,
                                unsigned int off_ac, unsigned int off_ad, unsigned int off_bc, unsigned int off_bd,
                                bool Tac, bool Tad, bool Tbc, bool Tbd);
/* Serving: K-builds of many clients (independent SCF runs) through one
   device-resident request loop.  A client is a handle (libgint_init done)
   whose call stream was recorded as a resident plan (plan_begin / plan_end);
   add_client takes it over (its own stream from then on; do not use the
   handle otherwise while the server runs) and keeps its density and
   exchange matrix (P_size doubles each) on the device between requests.
   submit queues one K-build: P is read and K overwritten (host arrays of
   P_size doubles, kept valid until wait returns); requests run in arrival
   order on a worker thread as clients become idle and fit the device
   budget (plan records, work buffers, tables, P, K of the resident clients
   <= max_mem_MiB; idle clients are evicted least recently used first, their
   plans kept on the host).  `slots` asks for several workers, whose
   requests would overlap on the GPU; one runs (with four, long runs stopped
   on the GPU, cause not found).  wait blocks until a request is done: out[0] latency
   (submit to K on the host), [1] queueing, [2] execution, seconds; returns 0,
   or -1 when the client can never fit the budget.  stats: out[0] requests
   done, [1] p50 and [2] p99 latency, [3] mean latency, [4] first arrival to
   last completion (s), [5] peak resident bytes, [6] evictions, [7]
   admissions.  destroy waits for the workers; the client handles remain the
   caller's. */
void *libgint_amd_server_create(int max_mem_MiB, int slots);
int libgint_amd_server_add_client(void *server, void *handle, int P_size);
int64_t libgint_amd_server_submit(void *server, int client, const double *P, double *K);
int libgint_amd_server_wait(void *server, int64_t request, double out[3]);
void libgint_amd_server_stats(void *server, double out[8]);
void libgint_amd_server_destroy(void *server);
/* device bytes the budget counts for a client (measured after its first
   request, estimated before), and a new budget (MiB) */
int64_t libgint_amd_server_client_bytes(void *server, int client);
void libgint_amd_server_set_budget(void *server, int max_mem_MiB);
/* the loop's state now: out[0] requests queued, [1] running, [2] done,
   [3] refused, [4] resident clients, [5] busy clients, [6] resident bytes,
   [7] budget bytes (for monitoring a serving process) */
void libgint_amd_server_state(void *server, int64_t out[8]);
/* reads CP2K's truncated-Coulomb coefficient file (the T_C_G_DATA
   t_c_g.dat) as reference src/t_c_g0_n.cpp:29-51 (read_c0) does: 207
   patches of 22 x 105 white-space separated numbers, patch after patch.
   Keeps the first (nder+1)*105 numbers of each patch in C0, row stride
   *ld_C0 = (nder+1)*105, ready for libgint_set_Potential_Truncated (host
   only, no GPU call).  Returns the number of doubles stored, 207 * *ld_C0,
   or -1 the file cannot be opened, -2 nder outside 0..21, -3 capacity below
   207*(nder+1)*105, -4 the file ends early or holds a token that is not a
   number (the reference aborts on these) */
int libgint_amd_read_c0(const char *path, int nder, double *C0, int capacity, int *ld_C0);
/* writes a full table (ld_C0 = 2310, orders 0..21) in that format, one
   patch per line, 17 significant digits so that reading it back is exact;
   returns 0, -1 on an I/O error, -2 on a bad argument */
int libgint_amd_write_c0(const char *path, const double *C0, int ld_C0);
/* D (16 x 16, row-major) = A (16 x 4) B (4 x 16) on the matrix-core tile
   the contraction kernel uses (prec 0: v_mfma_f64_16x16x4f64, prec 1:
   v_mfma_f32_16x16x4f32, inputs rounded to float), with that kernel's lane
   layout: a check of the layout against a host product; 0 or -1 */
int libgint_amd_mfma_selftest(const double *A, const double *B, double *D, int prec);
/* evaluates the screening arithmetic on the GPU for a bitwise comparison
   with the host: out[4i..4i+3] = exp(x), sqrt|x|, 1/x, x*y/(x+y) */
int libgint_amd_math_selftest(const double *x, const double *y, int n, double *out);
/* host evaluation of the same functions (same code) */
void libgint_amd_math_host(const double *x, const double *y, int n, double *out);

#ifdef __cplusplus
}
#endif
#endif
