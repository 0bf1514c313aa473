// Contraction of primitive blocks on the matrix cores (reference
// src/compute_ECO.cpp:12-214).  For one (n1,n2) row of a set quartet and one
// class (la lb|lc ld) whose block set has many columns and whose shells have
// many contractions (an spdf centre: 2-4 per shell, 16-256 per class):
//
//    C[q][j] = sum_p W[p][q] A[p][src(j)]
//
// A[p][.] the row's cell-summed primitive (e0|f0) blocks (k_fm_r), src the
// class's columns among the type's, W[p][q] = Ka[ia][ipa] Kb[ib][ipb]
// Kc[ic][ipc] Kd[id][ipd] the coefficient product of contraction
// q = (ia, ib, ic, id) (row-major, as the add_qrtt calls of the class).  C
// goes to the contraction buffer, row (r - first row of the chunk), where
// the per-class kernel (k_gqs) takes it in place of its own contraction loop.
//
// A tile of v_mfma_f64_16x16x4f64 is 16 columns x 16 contractions over 4
// primitive quartets.  A workgroup of 4 wavefronts takes 4 column tiles of
// one class and up to 64 contractions (one to four 16-wide tiles per
// wavefront, all accumulated in registers) and walks the row's primitive
// quartets 16 at a time: their coefficient products are formed once in LDS
// and serve the 4 wavefronts; each lane loads one block element per 4
// primitive quartets (16 consecutive columns per primitive quartet), the
// next batch's while the current one runs on the matrix cores.
//
// A row of a large type (an spdf quartet: thousands of primitive quartets x
// 7056 columns) is far larger than the caches, and the classes share its
// columns, so the items are ordered by the first column they read and an
// XCD takes a contiguous run of them: items that read the same columns run
// together, through one L2.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace gint {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Lane layout of v_mfma_f64_16x16x4f64 (D = A B + C, A 16x4, B 4x16):
//   lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
//   element r of lane l's D is D[4 r + l / 16][l % 16].
// libgint_amd_mfma_selftest checks it against a host product (rows of D
// are the rows of A as the lanes load them).
__device__ __forceinline__ dbl4 mfma_f64_16x16x4(double a, double b, dbl4 c)
{
   return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ __launch_bounds__(64) void k_mfma_selftest(const double *A, const double *B, double *D)
{
   const int l = threadIdx.x;
   dbl4 c = {0.0, 0.0, 0.0, 0.0};
   c = mfma_f64_16x16x4(A[(l % 16) * 4 + l / 16], B[(l / 16) * 16 + l % 16], c);
#pragma unroll
   for (int r = 0; r < 4; r++) D[(4 * r + l / 16) * 16 + l % 16] = c[r];
}

typedef float flt4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x4f32, the single-precision tile of the mixed-precision
// contraction; lane layout checked by libgint_amd_mfma_selftest (prec 1)
__device__ __forceinline__ flt4 mfma_f32_16x16x4(float a, float b, flt4 c)
{
   return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// element r of lane l's D in the single-precision tile: D[ECO32_ROW(l, r)][l % 16]
#define ECO32_ROW(l, r) (4 * ((l) / 16) + (r))

__global__ __launch_bounds__(64) void k_mfma32_selftest(const double *A, const double *B, double *D)
{
   const int l = threadIdx.x;
   flt4 c = {0.f, 0.f, 0.f, 0.f};
   c = mfma_f32_16x16x4((float)A[(l % 16) * 4 + l / 16], (float)B[(l / 16) * 16 + l % 16], c);
#pragma unroll
   for (int r = 0; r < 4; r++) D[ECO32_ROW(l, r) * 16 + l % 16] = (double)c[r];
}

int mfma_selftest(const double *A, const double *B, double *D, int prec)
{
   double *d = nullptr;
   int rc = 0;
   if (hipMalloc(&d, sizeof(double) * (64 + 64 + 256)) != hipSuccess) return -1;
   if (hipMemcpy(d, A, sizeof(double) * 64, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(d + 64, B, sizeof(double) * 64, hipMemcpyHostToDevice) != hipSuccess)
      rc = -1;
   if (!rc) {
      if (prec)
         hipLaunchKernelGGL(k_mfma32_selftest, dim3(1), dim3(64), 0, 0, d, d + 64, d + 128);
      else
         hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, 0, d, d + 64, d + 128);
      if (hipGetLastError() != hipSuccess) rc = -1;
   }
   if (!rc && hipMemcpy(D, d + 128, sizeof(double) * 256, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
   (void)hipFree(d);
   return rc;
}

constexpr int ECO_PB = 16;   // primitive quartets per LDS batch of coefficient products
constexpr int ECO_QB = 64;   // contractions per work item
constexpr int ECO_JB = 64;   // columns per work item (4 wavefronts x 16)

// one batch of ECO_PB primitive quartets into the wavefront's NQT tiles of
// 16 contractions: a[kk] the lane's block element of primitive quartet
// 4 kk + lane / 16, Wl = &W[lane / 16][lane % 16].  Single precision (f32)
// sums the batch in a single-precision tile and adds it to the double
// accumulators.
template <int NQT>
__device__ __forceinline__ void eco_batch(const double *a, const double *Wl, dbl4 *acc, bool f32)
{
   if (f32) {
      flt4 s32[NQT];
#pragma unroll
      for (int t = 0; t < NQT; t++) s32[t] = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < ECO_PB / 4; kk++)
#pragma unroll
         for (int t = 0; t < NQT; t++) s32[t] = mfma_f32_16x16x4((float)a[kk], (float)Wl[4 * kk * ECO_QB + 16 * t], s32[t]);
#pragma unroll
      for (int t = 0; t < NQT; t++)
#pragma unroll
         for (int rr = 0; rr < 4; rr++) acc[t][rr] += (double)s32[t][rr];
   } else {
#pragma unroll
      for (int kk = 0; kk < ECO_PB / 4; kk++)
#pragma unroll
         for (int t = 0; t < NQT; t++) acc[t] = mfma_f64_16x16x4(a[kk], Wl[4 * kk * ECO_QB + 16 * t], acc[t]);
   }
}

// three wavefronts per SIMD: the 4 x 16 double accumulators, two batches of
// block elements and the tile operands fit in 168 registers without spills
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_eco(const TypePlan *__restrict__ tpp, BatchDev b, EcoArgs e)
{
   const TypePlan &tp = *tpp;
   __shared__ double co[4][MAXNL][MAXNP];   // the class's four shells' coefficients
   __shared__ double W[ECO_PB][ECO_QB];     // coefficient products of the batch
   const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
   const uint32_t nrow = b.r1 - b.r0;
   const uint64_t nitems = (uint64_t)nrow * (uint64_t)e.n_items;
   // XCD-aware order: workgroups go to the 8 XCDs round robin, so XCD x
   // takes the x-th eighth of each grid's worth of items, a contiguous run
   // that reads the same columns (the items are ordered by column) through
   // its own L2
   const uint32_t wg = (gridDim.x & 7) ? blockIdx.x : (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
   for (uint64_t it = wg; it < nitems; it += gridDim.x) {
      const uint32_t r = b.r0 + (uint32_t)(it / (uint64_t)e.n_items);
      const EcoItem item = e.items[it % (uint64_t)e.n_items];
      const int k = item.k;
      const RowRec row = b.rows[r];
      const GroupRec &g = b.groups[row.group];
      const int sh = tp.cls_sh[k];
      const int shv[4] = {sh & 15, (sh >> 4) & 15, (sh >> 8) & 15, (sh >> 12) & 15};
      const int nlb = tp.nl[1][shv[1]], nlc = tp.nl[2][shv[2]], nld = tp.nl[3][shv[3]];
      const int nq = tp.cls_nl[k], cols = tp.cls_ncp[k];
      const int ncpt = tp.ncp;
      __syncthreads();   // the previous item no longer reads co, W
      for (int i = tid; i < 4 * MAXNL * MAXNP; i += 256) {
         const int X = i / (MAXNL * MAXNP), inl = (i / MAXNP) % MAXNL, ip = i % MAXNP;
         const int np = ((g.np4 >> (4 * X)) & 15) + 1;
         co[X][inl][ip] = (inl < tp.nl[X][shv[X]] && ip < np) ? b.data[g.Kof[X][shv[X]] + inl * np + ip] : 0.0;
      }
      // the wavefront's 16 columns and up to 4 tiles of 16 contractions
      const int jt = item.jt0 + wv;
      const int j = jt * 16 + (lane & 15);
      const bool jv = j < cols;
      const bool tv = jt * 16 < cols;
      const int src = jv ? e.srcmaps[e.src_off[k] + j] : 0;
      const int q0 = item.qb * ECO_QB;
      const int nqt = min(4, (nq - q0 + 15) / 16);
      // this thread's coefficient products: contraction q0 + (tid & 63) for
      // primitive quartets tid / 64 + 4 m of each batch
      const int qq = tid & 63, q = q0 + qq;
      const bool qv = q < nq;
      const int ia = q / (nlb * nlc * nld), ib = (q / (nlc * nld)) % nlb, ic = (q / nld) % nlc, id = q % nld;
      // mixed precision: single-precision tiles for a set quartet whose
      // bound is small (each batch of 16 primitive quartets summed in single
      // precision, the batches in double precision)
      const bool f32 = b.gflag && b.gflag[row.group];
      dbl4 acc[4];
#pragma unroll
      for (int t = 0; t < 4; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
      const double *A0 = b.ac0 + (size_t)(row.prm_off - b.p0) * ncpt + src;
      const uint32_t *IP = b.prm_ip + row.prm_off;
      const int n = (int)row.n_prm;
      // the next batch's block elements and primitive indices, loaded while
      // the current batch runs
      double an[ECO_PB / 4];
      uint32_t ipn[ECO_PB / 4];
#pragma unroll
      for (int kk = 0; kk < ECO_PB / 4; kk++) {
         const int p = 4 * kk + (lane >> 4);
         an[kk] = (jv && p < n) ? A0[(size_t)p * ncpt] : 0.0;
         const int pp = (tid >> 6) + 4 * kk;
         ipn[kk] = pp < n ? IP[pp] : 0u;
      }
      for (int pb = 0; pb < n; pb += ECO_PB) {
         double a[ECO_PB / 4];
         uint32_t ipc[ECO_PB / 4];
#pragma unroll
         for (int kk = 0; kk < ECO_PB / 4; kk++) {
            a[kk] = an[kk];
            ipc[kk] = ipn[kk];
         }
         __syncthreads();   // the previous batch's W is no longer read
#pragma unroll
         for (int m = 0; m < ECO_PB / 4; m++) {
            const int pp = (tid >> 6) + 4 * m;
            const uint32_t ip = ipc[m];
            W[pp][qq] = (qv && pb + pp < n) ? co[0][ia][ip & 15] * co[1][ib][(ip >> 4) & 15] *
                                                 co[2][ic][(ip >> 8) & 15] * co[3][id][(ip >> 12) & 15]
                                           : 0.0;
         }
         __syncthreads();
         if (pb + ECO_PB < n) {
#pragma unroll
            for (int kk = 0; kk < ECO_PB / 4; kk++) {
               const int p = pb + ECO_PB + 4 * kk + (lane >> 4);
               an[kk] = (jv && p < n) ? A0[(size_t)p * ncpt] : 0.0;
               const int pp = pb + ECO_PB + (tid >> 6) + 4 * kk;
               ipn[kk] = pp < n ? IP[pp] : 0u;
            }
         }
         if (tv) {
            const double *Wl = &W[lane >> 4][lane & 15];
            switch (nqt) {
               case 1: eco_batch<1>(a, Wl, acc, f32); break;
               case 2: eco_batch<2>(a, Wl, acc, f32); break;
               case 3: eco_batch<3>(a, Wl, acc, f32); break;
               default: eco_batch<4>(a, Wl, acc, f32); break;
            }
         }
      }
      if (!tv) continue;
      double *C = b.eco + (size_t)(r - b.r0) * e.row_doubles + e.eco_off[k];
#pragma unroll
      for (int t = 0; t < 4; t++) {
         const int qo = q0 + 16 * t + (lane & 15);
         if (t >= nqt || qo >= nq) continue;
#pragma unroll
         for (int rr = 0; rr < 4; rr++) {
            // the two tiles place element rr of a lane in different rows
            const int jj = jt * 16 + (f32 ? ECO32_ROW(lane, rr) : 4 * rr + (lane >> 4));
            if (jj < cols) C[(size_t)qo * cols + jj] = acc[t][rr];
         }
      }
   }
}

int launch_eco(const TypePlan &tp, const BatchDev &b, const EcoArgs &e, void *stream)
{
   if (e.n_items <= 0 || b.r1 <= b.r0) return 0;
   (void)tp;
   const uint64_t nitems = (uint64_t)(b.r1 - b.r0) * (uint64_t)e.n_items;
   unsigned grid = (unsigned)std::min<uint64_t>(nitems, 65536);
   if (grid >= 64) grid &= ~7u;   // a multiple of the 8 XCDs (the kernel's item order)
   hipLaunchKernelGGL(k_eco, dim3(grid), dim3(256), 0, (hipStream_t)stream, b.tp, b, e);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace gint
