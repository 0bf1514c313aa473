// Host-side tabulation of the Boys function F_n(T) on T = 0, 0.1, ..., 12 for
// n = 0..27, the table the T < 12 branch of gint_boys expands around.
// Reference src/fgamma.cpp:53-123 (series in ratios of modified spherical
// Bessel functions, asymptotic form above T = 50), called with nmax = 21,
// tdelta = 0.1, tmax = 12 at src/libGint.cpp:488-497.
#include <cfloat>
#include <cmath>

namespace gint {

static void boys_series(int nmax, double T, double *f)
{
   const int kmax = 50;
   double r[kmax + 10];
   if (T < 1.e-13) {
      for (int n = 0; n <= nmax; n++) f[n] = 1.0 / (2. * n + 1.);
   } else if (T <= 50.0) {
      r[kmax + 9] = 0.0;
      for (int j = kmax + 9; j >= 1; j--) r[j - 1] = -T / ((4. * j + 2.) - T * r[j]);
      const double factor = 2.0 * std::sinh(0.5 * T) * std::exp(-0.5 * T) / T;
      for (int n = 0; n <= nmax; n++) {
         double sumtot = factor / (2. * n + 1.);
         double term = 1.0;
         for (int k = 1; k <= kmax; k++) {
            term = term * (2. * n - 2. * k + 1.) / (2. * n + 2. * k + 1);
            double p = 1.0;
            for (int j = 1; j <= k; j++) p = p * r[j - 1];
            const double sumterm = factor * term * p * (2. * k + 1.) / (2. * n + 1.);
            if (std::fabs(sumterm) < DBL_EPSILON) break;
            sumtot = sumtot + sumterm;
         }
         f[n] = sumtot;
      }
   } else {
      f[0] = 0.5 * std::sqrt(M_PI / T);
      const double expt = std::exp(-T);
      for (int n = 1; n <= nmax; n++) f[n] = 0.5 * ((2. * n - 1.) * f[n - 1] - expt) / T;
   }
}

void boys_table(double *ftable)
{
   const int n = 21 + 6;
   const double tdelta = 0.1;
   const int itabmax = (int)((12. - 0.) / tdelta);
   for (int itab = 0; itab <= itabmax; itab++) boys_series(n, (double)itab * tdelta, &ftable[itab * (n + 1)]);
}

}  // namespace gint
