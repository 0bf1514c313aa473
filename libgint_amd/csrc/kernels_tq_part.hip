// One of GINT_TQ_NPART translation units of the templated per-class quartet
// kernels (kernels_tq.hpp): built once per part with -DGINT_TQ_PART=<n>, so
// that the 58 instantiations compile in parallel.
#include "kernels_tq.hpp"

#ifndef GINT_TQ_PART
#error "build with -DGINT_TQ_PART=<part>"
#endif

namespace gint {
namespace tq {

template <int la, int lb, int lc, int ld>
int launch_one(int cls_k, const TypePlan &tp, const DevTables &t, const BatchDev &b, const KTargets &kt,
               const int *srcmap, hipStream_t st)
{
   return launch_wave<la, lb, lc, ld>(cls_k, tp, t, b, kt, srcmap, st);
}

#define GINT_TQ_INST(a, b, c, d)                                                                                   \
   template int launch_one<a, b, c, d>(int, const TypePlan &, const DevTables &, const BatchDev &, const KTargets &, \
                                       const int *, hipStream_t);
#if GINT_TQ_PART == 0
GINT_TQ_CLASSES_0(GINT_TQ_INST)
#elif GINT_TQ_PART == 1
GINT_TQ_CLASSES_1(GINT_TQ_INST)
#elif GINT_TQ_PART == 2
GINT_TQ_CLASSES_2(GINT_TQ_INST)
#elif GINT_TQ_PART == 3
GINT_TQ_CLASSES_3(GINT_TQ_INST)
#elif GINT_TQ_PART == 4
GINT_TQ_CLASSES_4(GINT_TQ_INST)
#elif GINT_TQ_PART == 5
GINT_TQ_CLASSES_5(GINT_TQ_INST)
#elif GINT_TQ_PART == 6
GINT_TQ_CLASSES_6(GINT_TQ_INST)
#elif GINT_TQ_PART == 7
GINT_TQ_CLASSES_7(GINT_TQ_INST)
#endif

}  // namespace tq
}  // namespace gint
