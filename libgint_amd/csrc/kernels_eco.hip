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
// primitive quartets (16 consecutive columns per primitive quartet).
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

__global__ __launch_bounds__(256) void k_eco(const TypePlan *__restrict__ tpp, BatchDev b, EcoArgs e)
{
   const TypePlan &tp = *tpp;
   __shared__ double co[4][MAXNL][MAXNP];   // the class's four shells' coefficients
   __shared__ double W[ECO_PB][ECO_QB];     // coefficient products of the batch
   __shared__ uint32_t ipl[ECO_PB];
   const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
   const uint32_t nrow = b.r1 - b.r0;
   const uint64_t nitems = (uint64_t)nrow * (uint64_t)e.n_items;
   for (uint64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
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
      __syncthreads();   // the previous item no longer reads co, W, ipl
      for (int i = tid; i < 4 * MAXNL * MAXNP; i += 256) {
         const int X = i / (MAXNL * MAXNP), inl = (i / MAXNP) % MAXNL, ip = i % MAXNP;
         const int np = ((g.np4 >> (4 * X)) & 15) + 1;
         co[X][inl][ip] = (inl < tp.nl[X][shv[X]] && ip < np) ? b.data[g.Kof[X][shv[X]] + inl * np + ip] : 0.0;
      }
      // the wavefront's 16 columns and up to 4 tiles of 16 contractions
      const int jt = item.jt0 + wv;
      const int j = jt * 16 + (lane & 15);
      const bool jv = j < cols;
      const int src = jv ? e.srcmaps[e.src_off[k] + j] : 0;
      const int q0 = item.qb * ECO_QB;
      const int nqt = min(4, (nq - q0 + 15) / 16);
      // mixed precision: single-precision tiles for a set quartet whose
      // bound is small (each batch of 16 primitive quartets summed in single
      // precision, the batches in double precision)
      const bool f32 = b.gflag && b.gflag[row.group];
      dbl4 acc[4];
#pragma unroll
      for (int t = 0; t < 4; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
      const double *A0 = b.ac0 + (size_t)(row.prm_off - b.p0) * ncpt;
      const uint32_t *IP = b.prm_ip + row.prm_off;
      const int n = (int)row.n_prm;
      for (int pb = 0; pb < n; pb += ECO_PB) {
         __syncthreads();
         if (tid < ECO_PB) ipl[tid] = pb + tid < n ? IP[pb + tid] : 0u;
         __syncthreads();
         for (int i = tid; i < ECO_PB * ECO_QB; i += 256) {
            const int pp = i / ECO_QB, qq = i % ECO_QB, q = q0 + qq;
            double w = 0.0;
            if (pb + pp < n && q < nq) {
               const uint32_t ip = ipl[pp];
               const int ia = q / (nlb * nlc * nld), ib = (q / (nlc * nld)) % nlb, ic = (q / nld) % nlc, id = q % nld;
               w = co[0][ia][ip & 15] * co[1][ib][(ip >> 4) & 15] * co[2][ic][(ip >> 8) & 15] *
                   co[3][id][(ip >> 12) & 15];
            }
            W[pp][qq] = w;
         }
         __syncthreads();
         if (jt * 16 < cols) {
            double a[ECO_PB / 4];
#pragma unroll
            for (int kk = 0; kk < ECO_PB / 4; kk++) {
               const int p = pb + 4 * kk + (lane >> 4);
               a[kk] = (jv && p < n) ? A0[(size_t)p * ncpt + src] : 0.0;
            }
            if (f32) {
               flt4 s32[4];
#pragma unroll
               for (int t = 0; t < 4; t++) s32[t] = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
               for (int kk = 0; kk < ECO_PB / 4; kk++) {
                  const double *Wk = &W[4 * kk + (lane >> 4)][lane & 15];
#pragma unroll
                  for (int t = 0; t < 4; t++)
                     if (t < nqt) s32[t] = mfma_f32_16x16x4((float)a[kk], (float)Wk[16 * t], s32[t]);
               }
#pragma unroll
               for (int t = 0; t < 4; t++)
#pragma unroll
                  for (int rr = 0; rr < 4; rr++) acc[t][rr] += (double)s32[t][rr];
            } else {
#pragma unroll
               for (int kk = 0; kk < ECO_PB / 4; kk++) {
                  const double *Wk = &W[4 * kk + (lane >> 4)][lane & 15];
#pragma unroll
                  for (int t = 0; t < 4; t++)
                     if (t < nqt) acc[t] = mfma_f64_16x16x4(a[kk], Wk[16 * t], acc[t]);
               }
            }
         }
      }
      if (jt * 16 >= cols) continue;
      double *C = b.eco + (size_t)(r - b.r0) * e.row_doubles + e.eco_off[k];
#pragma unroll
      for (int t = 0; t < 4; t++) {
         const int q = q0 + 16 * t + (lane & 15);
         if (t >= nqt || q >= nq) continue;
#pragma unroll
         for (int rr = 0; rr < 4; rr++) {
            // the two tiles place element rr of a lane in different rows
            const int jj = jt * 16 + (f32 ? ECO32_ROW(lane, rr) : 4 * rr + (lane >> 4));
            if (jj < cols) C[(size_t)q * cols + jj] = acc[t][rr];
         }
      }
   }
}

int launch_eco(const TypePlan &tp, const BatchDev &b, const EcoArgs &e, void *stream)
{
   if (e.n_items <= 0 || b.r1 <= b.r0) return 0;
   (void)tp;
   const uint64_t nitems = (uint64_t)(b.r1 - b.r0) * (uint64_t)e.n_items;
   const unsigned grid = (unsigned)std::min<uint64_t>(nitems, 65536);
   hipLaunchKernelGGL(k_eco, dim3(grid), dim3(256), 0, (hipStream_t)stream, b.tp, b, e);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace gint
