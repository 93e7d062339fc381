"""Landau-damping diagnostics of the Vlasov-Poisson runs (host side).

The reference writes the electric energy per component every diagnostics
tick (examples/vlasov_poisson/include/application.h:630-660,
diagnostics.h:88-141); the damping rate is read off its decay.  These
helpers turn an energy trace into a rate and give the rate of the linear
dispersion relation to compare against.
"""

import numpy as np


def fit_damping_rate(t, energy, t0=0.5):
    """gamma from a least-squares line through log(E) at the local maxima of
    E(t) for t >= t0 (the field energy E ~ |E|^2 decays like exp(2 gamma t));
    returns (gamma, number of maxima used), gamma NaN below three maxima."""
    t, e = np.asarray(t, dtype=float), np.asarray(energy, dtype=float)
    idx = [i for i in range(1, len(e) - 1)
           if e[i] >= e[i - 1] and e[i] > e[i + 1] and t[i] >= t0]
    if len(idx) < 3:
        return float("nan"), len(idx)
    slope = np.polyfit(t[idx], np.log(e[idx]), 1)[0]
    return float(slope / 2.0), len(idx)


def landau_theory_rate(k):
    """(omega, gamma) of the Langmuir root of the linear Landau dispersion
    relation 1 + (1 + z Z(z)) / k^2 = 0, z = omega / (sqrt(2) k), for a
    Maxwellian of unit thermal speed (Z: the plasma dispersion function,
    i sqrt(pi) w(z) with the Faddeeva function w)."""
    from scipy.optimize import fsolve
    from scipy.special import wofz

    def eps(w):
        z = w / (np.sqrt(2.0) * k)
        zz = 1j * np.sqrt(np.pi) * wofz(z)
        return 1.0 + (1.0 + z * zz) / k ** 2

    sol = fsolve(lambda x: [eps(x[0] + 1j * x[1]).real,
                            eps(x[0] + 1j * x[1]).imag], [1.7, -0.4])
    return float(sol[0]), float(sol[1])
