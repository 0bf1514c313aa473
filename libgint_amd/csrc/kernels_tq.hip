// Dispatcher of the templated per-class quartet kernels (kernels_tq.hpp):
// which classes have one, and the launch of class k of a type.
#include "kernels_tq.hpp"

namespace gint {

bool tq_has(int code)
{
   switch (code) {
#define GINT_TQ_CASE(a, b, c, d) case class_code(a, b, c, d):
      GINT_TQ_CLASSES(GINT_TQ_CASE)
      return true;
#undef GINT_TQ_CASE
      default:
         return false;
   }
}

int launch_tq(int cls_k, const TypePlan &tp, const DevTables &t, const BatchDev &b, const KTargets &kt,
              const int *srcmap, void *stream)
{
   if (b.g1 <= b.g0) return 0;
   hipStream_t st = (hipStream_t)stream;
   switch (tp.cls_code[cls_k]) {
#define GINT_TQ_CASE(a, b_, c, d) \
   case class_code(a, b_, c, d): return tq::launch_one<a, b_, c, d>(cls_k, tp, t, b, kt, srcmap, st);
      GINT_TQ_CLASSES(GINT_TQ_CASE)
#undef GINT_TQ_CASE
      default:
         return -5;
   }
}

}  // namespace gint
