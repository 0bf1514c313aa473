// The truncated-Coulomb coefficient file CP2K names with T_C_G_DATA
// (t_c_g.dat), as the reference reads it (src/t_c_g0_n.cpp:29-51): 207
// patches, each holding the 105 coefficients of a degree-13 triangular
// Chebyshev expansion for every order 0..21 (2310 numbers), patch after
// patch, separated by white space.  A reader asked for orders 0..nder keeps
// the first (nder+1)*105 numbers of each patch and skips the rest, which
// gives the row layout libgint_set_Potential_Truncated takes (row stride
// ld_C0 = (nder+1)*105).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/libGint.h"

namespace {

constexpr int kPatches = 207;
constexpr int kPerOrder = 105;   // (13+1)(13+2)/2
constexpr int kOrders = 22;      // orders 0..21 in the file
constexpr int kPerPatch = kOrders * kPerOrder;

// one white-space separated number; false at the end of the file or on a
// token that is not a number
bool next_number(FILE *f, double *v)
{
   char tok[64];
   if (std::fscanf(f, "%63s", tok) != 1) return false;
   char *end = nullptr;
   errno = 0;
   const double x = std::strtod(tok, &end);
   if (end == tok || *end != '\0' || errno == ERANGE) return false;
   *v = x;
   return true;
}

}  // namespace

extern "C" {

int libgint_amd_read_c0(const char *path, int nder, double *C0, int capacity, int *ld_C0)
{
   if (nder < 0 || nder >= kOrders) return -2;
   const int ld = (nder + 1) * kPerOrder;
   if (!C0 || capacity < kPatches * ld) return -3;
   FILE *f = path ? std::fopen(path, "r") : nullptr;
   if (!f) return -1;
   for (int p = 0; p < kPatches; p++)
      for (int j = 0; j < kPerPatch; j++) {
         double v;
         if (!next_number(f, &v)) {
            std::fclose(f);
            return -4;
         }
         if (j < ld) C0[p * ld + j] = v;
      }
   std::fclose(f);
   if (ld_C0) *ld_C0 = ld;
   return kPatches * ld;
}

int libgint_amd_write_c0(const char *path, const double *C0, int ld_C0)
{
   if (!C0 || ld_C0 != kPerPatch) return -2;
   FILE *f = path ? std::fopen(path, "w") : nullptr;
   if (!f) return -1;
   bool ok = true;
   for (int p = 0; p < kPatches && ok; p++) {
      // one patch per line; 17 significant digits read back exactly
      for (int j = 0; j < kPerPatch && ok; j++)
         ok = std::fprintf(f, j ? " %.16E" : "%.16E", C0[p * kPerPatch + j]) > 0;
      ok = ok && std::fputc('\n', f) != EOF;
   }
   ok = (std::fclose(f) == 0) && ok;
   return ok ? 0 : -1;
}

}  // extern "C"
