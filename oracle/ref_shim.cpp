// Test-only shim: compiles the reference's self-contained plain-C pieces in
// place (the patch decision trees, the patch bias/width table and the
// cartesian-to-spherical literals) into oracle/_ref/libref.so so the tests
// can compare this build's restated tables against them.  Built only where
// the reference checkout exists; never linked into the product.
#include "locate_patch_low_R.cpp"
#include "locate_patch_high_R.cpp"
#include "BW_by_patch.h"
#include "c2s.h"

extern "C" {
int ref_locate_patch_low_R(double x1, double x2) { return locate_patch_low_R(x1, x2); }
int ref_locate_patch_high_R(double x1, double x2) { return locate_patch_high_R(x1, x2); }
double ref_bw_by_patch(int i) { return BW_by_patch[i]; }
double ref_c2s(int i) { return c2s[i]; }
int ref_c2s_ptr(int l) { return c2s_ptr[l]; }
}
