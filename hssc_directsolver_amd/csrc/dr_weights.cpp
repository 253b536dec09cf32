// Order-4 Duan-Rokhlin centre weight for the 2D Helmholtz kernel H0(k r) on the lattice hZ^2.
//
// Replaces Common/DR_stencil.m:64-107 (DVEC(1) from the Fortran side-car
// Common/FORTRAN_get_DR_weights.f:85-490).  The weight w makes
//   sum_{j != 0} h^2 H0(k|x_j|) phi(x_j) + w phi(0)
// a 4th-order rule for int H0(k|x|) phi(x) dx.  With Y0(kr) = (2/pi) log(r) J0(kr) + smooth,
// smooth terms only need the centre point (h^2 g(0)), and log(r) r^(2m) needs the centre
// weight h^(2+2m) (Z'(-2m) - log(h) Z(-2m)) from the generalized Euler-Maclaurin formula,
// Z(s) = sum_{j != 0} |j|^-s = 4 zeta(s/2) beta(s/2) the Epstein zeta function of Z^2
// (Z(0) = -1, Z(-2m) = 0 for m >= 1).  Summing over the Taylor series of J0:
//   corr0 = h^2 + i (2h^2/pi) [log(hk/2) + gamma + Z'(0)
//                              + sum_{m>=1} (-1)^m (hk/2)^(2m) / (m!)^2 Z'(-2m)].

#include <cmath>

#include "internal.h"

namespace {

// Z'(-2m), m = 0..15, from Z'(s) = 2 [zeta'(s/2) beta(s/2) + zeta(s/2) beta'(s/2)]
// evaluated in 50-digit arithmetic.
constexpr double kEpsteinDz[16] = {
    -1.31053292591150951825,  -0.0971869680102729267128, 0.0304484570583932707803,
    -0.0255159834104026340728, 0.0399190572513431214035, -0.0994580498291840983112,
    0.359885307754472184488,  -1.77969261333065431128,  11.5178843500591125929,
    -94.4349799029697638983,  956.358808529189846422,   -11721.9645921458875723,
    171005.518184889246804,   -2927996.63911569921292,  58145170.1211665987512,
    -1325530701.7662366223};

constexpr double kEulerGamma = 0.57721566490153286061;
constexpr double kPi = 3.14159265358979323846;

}  // namespace

extern "C" int hssc_dr_corr0(double h, double k, double* re, double* im) {
  if (!re || !im) return hssc::fail(HSSC_ERR_INVALID_ARG, "null output");
  const double hk = h * k;
  // The side-car's tables cover hk in [1e-4, 1.01] (FORTRAN_get_DR_weights.f:130-143 returns
  // ier = 4 outside); the series below is accurate to 1e-16 for 0 < hk <= 2.
  if (!(h > 0) || !(hk > 0 && hk <= 2.0))
    return hssc::fail(HSSC_ERR_INVALID_ARG, "Duan-Rokhlin weights need h > 0 and 0 < h*k <= 2");
  const double x = 0.5 * hk, x2 = x * x;
  // series terms (-1)^m x^(2m) / (m!)^2 Z'(-2m), summed from the smallest upwards
  double term[16];
  double c = 1.0;
  term[0] = 0.0;
  for (int m = 1; m < 16; ++m) {
    c *= -x2 / (double(m) * double(m));
    term[m] = c * kEpsteinDz[m];
  }
  double s = 0.0;
  for (int m = 15; m >= 1; --m) s += term[m];
  s += std::log(x) + kEulerGamma + kEpsteinDz[0];
  *re = h * h;
  *im = 2.0 * h * h / kPi * s;
  return HSSC_OK;
}
