// gmc_ppm.hpp -- the piecewise-parabolic reconstruction of the FFSL fluxes
// (ffsl_mod.F90:1147-1205), shared by the horizontal advection kernels and
// the vertical flux of the column operators.
#pragma once

#include "gmc_internal.hpp"
#include "../../include/gmcore_adv.h"

namespace gmc {

// min / max of the limiters: one v_min_f64 / v_max_f64 each.  They may
// break a tie between +0 and -0 differently from the oracle's a < b ? a : b,
// but every result of the slope and PPM limiters passes through fabs (fsign)
// or a difference with a value that is itself the tie, so the limiters'
// outputs carry the same bits either way (finite inputs).
__device__ __forceinline__ double dmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }

// slope_simple / slope_mono / slope_pd (ffsl_mod.F90:1175-1205)
__device__ __forceinline__ double slope(int lim, double fm1, double f, double fp1) {
  const double df = (fp1 - fm1) * 0.5;
  if (lim == GMC_LIMITER_NONE) return df;
  if (lim == GMC_LIMITER_PD) return fsign(dmin(fabs(df), 2 * f), df);
  const double df_min = 2 * (f - dmin(dmin(fm1, f), fp1));
  const double df_max = 2 * (dmax(dmax(fm1, f), fp1) - f);
  return fsign(dmin(dmin(fabs(df), df_min), df_max), df);
}

// ppm (ffsl_mod.F90:1147-1173)
struct Ppm { double fl, df, f6; };
__device__ __forceinline__ Ppm ppm(int lim, double fm2, double fm1, double f, double fp1, double fp2) {
  const double dfl = slope(lim, fm2, fm1, f);
  const double df = slope(lim, fm1, f, fp1);
  const double dfr = slope(lim, f, fp1, fp2);
  // x / 6 through div_rn: the IEEE quotient in three operations
  double fl = 0.5 * (fm1 + f) + div_rn(dfl - df, 6.0, 1.0 / 6.0);
  double fr = 0.5 * (fp1 + f) + div_rn(df - dfr, 6.0, 1.0 / 6.0);
  fl = f - fsign(dmin(fabs(df), fabs(fl - f)), df);
  fr = f + fsign(dmin(fabs(df), fabs(fr - f)), df);
  return Ppm{fl, fr - fl, 6 * f - 3 * (fl + fr)};
}

}  // namespace gmc
