/*
 * Cartesian -> real spherical transformation coefficients, CP2K convention.
 *
 * c2s_l is a (2l+1) x ncart(l) row-major matrix; row m runs over the
 * spherical functions in CP2K order (m = -l..l), column k over cartesian
 * components in lexicographic order with descending x
 * (d: xx xy xz yy yz zz).  The contraction coefficients are normalised so
 * that a cartesian function with every exponent <= 1 has unit norm
 * (see libgint_set_Atom_L), and these coefficients supply the remaining
 * factors.
 *
 * The values must be the ones the reference engine applies so that the
 * K-build agrees with it: reference src/c2s.h:36-97 (l <= 2 exact,
 * l = 3 as 9-significant-digit decimals, which are kept as given).
 */
#ifndef LIBGINT_C2S_TABLE_H
#define LIBGINT_C2S_TABLE_H

#define GINT_C2S_SIZE (1 * 1 + 3 * 3 + 5 * 6 + 7 * 10)
/* start of the block for l = 0..3 */
#define GINT_C2S_OFF_0 0
#define GINT_C2S_OFF_1 1
#define GINT_C2S_OFF_2 10
#define GINT_C2S_OFF_3 40

/* 0.5/sqrt(3.0) and 1.0/sqrt(3.0) as evaluated in double arithmetic (one ulp
   above the correctly rounded 1/(2 sqrt 3) and 1/sqrt 3) */
#define GINT_INV2SQRT3 0x1.279a74590331dp-2
#define GINT_INVSQRT3  0x1.279a74590331dp-1

/* the coefficients as a list, for host arrays and device constants alike */
#define GINT_C2S_VALUES \
   /* l = 0 */ \
   1.0, \
   /* l = 1 : y z x */ \
   0.0, 1.0, 0.0, \
   0.0, 0.0, 1.0, \
   1.0, 0.0, 0.0, \
   /* l = 2 : columns xx xy xz yy yz zz */ \
   0.0, 1.0, 0.0, 0.0, 0.0, 0.0, \
   0.0, 0.0, 0.0, 0.0, 1.0, 0.0, \
   -GINT_INV2SQRT3, 0.0, 0.0, -GINT_INV2SQRT3, 0.0, GINT_INVSQRT3, \
   0.0, 0.0, 1.0, 0.0, 0.0, 0.0, \
   0.5, 0.0, 0.0, -0.5, 0.0, 0.0, \
   /* l = 3 : columns xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz */ \
   0.0, 0.612372436, 0.0, 0.0, 0.0, 0.0, -0.204124145, 0.0, 0.0, 0.0, \
   0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, \
   0.0, -0.158113883, 0.0, 0.0, 0.0, 0.0, -0.158113883, 0.0, 0.632455532, 0.0, \
   0.0, 0.0, -0.387298335, 0.0, 0.0, 0.0, 0.0, -0.387298335, 0.0, 0.25819889, \
   -0.158113883, 0.0, 0.0, -0.158113883, 0.0, 0.632455532, 0.0, 0.0, 0.0, 0.0, \
   0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 0.0, -0.5, 0.0, 0.0, \
   0.204124145, 0.0, 0.0, -0.612372436, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0

static const double gint_c2s[GINT_C2S_SIZE] = {GINT_C2S_VALUES};

static const int gint_c2s_off[4] = {GINT_C2S_OFF_0, GINT_C2S_OFF_1, GINT_C2S_OFF_2, GINT_C2S_OFF_3};

#endif
