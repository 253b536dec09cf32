"""Duan-Rokhlin diagonal correction corr0(h, k) for the 2D Helmholtz kernel (oracle).

Test infrastructure only: imported by tests, the smoke check and the CPU baseline.

Replaces Common/DR_stencil.m:10-25,64-107 for order = 4 (dr_weights = DVEC(1)), whose
value comes from the reference's Fortran side-car (Common/FORTRAN_get_DR_weights.f:85-490,
corr_hank: Lagrange interpolation of precomputed tables).  Here corr0 is computed from the
mathematics instead, in multiple-precision arithmetic:

On the lattice hZ^2 the punctured trapezoidal rule plus a centre weight w,

    sum_{j != 0} h^2 H0(k|x_j|) phi(x_j) + w phi(0)  ~  int H0(k|x|) phi(x) dx,

is corrected at order 4 by w = corr0.  Write H0 = J0(kr) + i Y0(kr) with
Y0(kr) = (2/pi) log(r) J0(kr) + (2/pi)(log(k/2) + gamma) J0(kr) + (2/pi) P(kr), where the
last two terms are smooth (P is an even power series vanishing at 0).  Smooth terms need
only the centre point h^2 g(0).  For log(r) r^(2m) the generalized Euler-Maclaurin formula
for the Epstein zeta function Z(s) = sum_{j != 0} |j|^-s = 4 zeta(s/2) beta(s/2) of Z^2
gives the centre weight h^(2+2m) (Z'(-2m) - log(h) Z(-2m)), with Z(0) = -1 and Z(-2m) = 0
for m >= 1.  Summing over the Taylor coefficients of J0(kr):

    corr0 = h^2 + i (2 h^2 / pi) [ log(hk/2) + gamma + Z'(0)
                                   + sum_{m>=1} (-1)^m (hk/2)^(2m) / (m!)^2 Z'(-2m) ].

Pinned against the reference's published known answer (INSTALL.txt:96-115: h = 1/32,
rk = 1 gives corr0 = 0.976562500000008e-3 - 0.304146129302758e-2 i).
"""

from __future__ import annotations

import mpmath as mp


def epstein_dz(m: int, dps: int = 40):
    """Z'(-2m) for the square lattice, from Z(s) = 4 zeta(s/2) beta(s/2)."""
    with mp.workdps(dps):
        q1, q3 = mp.mpf(1) / 4, mp.mpf(3) / 4

        def beta(s):
            return (mp.zeta(s, q1) - mp.zeta(s, q3)) / mp.power(4, s)

        def dbeta(s):
            return (mp.zeta(s, q1, 1) - mp.zeta(s, q3, 1)) / mp.power(4, s) - mp.log(4) * beta(s)

        u = mp.mpf(-m)
        return 2 * (mp.zeta(u, 1, 1) * beta(u) + mp.zeta(u) * dbeta(u))


def corr0(h: float, k: float, terms: int = 16, dps: int = 40) -> complex:
    """Order-4 Duan-Rokhlin centre weight (DR_stencil(h, rk=k, n, 4) -> dr_weights)."""
    with mp.workdps(dps):
        h_, k_ = mp.mpf(h), mp.mpf(k)
        x = h_ * k_ / 2
        s = mp.log(x) + mp.euler + epstein_dz(0, dps)
        for m in range(1, terms):
            s += (-1) ** m * x ** (2 * m) / mp.factorial(m) ** 2 * epstein_dz(m, dps)
        im = 2 * h_ * h_ / mp.pi * s
        return complex(float(h_ * h_), float(im))
