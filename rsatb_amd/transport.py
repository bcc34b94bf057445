"""Finite-temperature DC conductivity from a zero-temperature spectrum.

The spectrum sigma(E) comes from the device (the Greenwood or Bastin
moment route); what is left is a one-dimensional integral over at most a few
thousand energies per k-point, done here on the host, as the reference does
it on its host (rsatb/calc/fftcheb_transport.hpp:950-1040,
rsatb/calc/observables.hpp:792-854).
"""
from __future__ import annotations

import math

import numpy as np

KB_EV = 8.617333262e-5  # Boltzmann constant, eV/K


def fermi_dirac_energy(energy, mu: float, temp: float):
    """Occupation f(E) (rsatb/calc/fftcheb_transport.hpp:67-81): a step
    (1 at E <= mu) at temp <= 0, else 1 / (1 + exp((E - mu) / kT)) written so
    that the exponent is never positive."""
    e = np.asarray(energy, dtype=np.float64)
    if temp <= 0.0:
        return np.where(e <= mu, 1.0, 0.0)
    t = np.exp(-np.abs(e - mu) / (KB_EV * temp))
    return np.where(e >= mu, t / (1.0 + t), 1.0 / (1.0 + t))


def interpolate_spectrum_component(energy, sigma, mu: float):
    """Linear interpolation of every column of sigma (n_energy, n_comp) at mu,
    clamped to the end values outside the grid
    (rsatb/calc/fftcheb_transport.hpp:950-976)."""
    e = np.asarray(energy, dtype=np.float64)
    s = np.asarray(sigma, dtype=np.float64)
    if e.size == 0:
        return np.zeros(s.shape[1:] if s.ndim > 1 else ())
    if mu <= e[0]:
        return s[0].copy()
    if mu >= e[-1]:
        return s[-1].copy()
    hi = int(np.searchsorted(e, mu, side="left"))
    lo = hi - 1
    t = (mu - e[lo]) / (e[hi] - e[lo])
    return (1.0 - t) * s[lo] + t * s[hi]


def integrate_dc_spectrum(sigma_spectrum, energy, mu: float, temp: float):
    """sigma(mu, T) = integral of -df/dE sigma(E) dE with sigma taken piecewise
    linear between the grid energies (rsatb/calc/fftcheb_transport.hpp:995-1040).

    On a segment [e0, e1] with sigma = s0 + m (E - e0) the integral is exact:
    s0 (f(e0) - f(e1)) + m (F(e1) - F(e0) - f(e1) (e1 - e0)), where
    F(E) = E - kT softplus((E - mu) / kT) is an antiderivative of f. At
    temp <= 0 the result is the spectrum interpolated at mu. sigma_spectrum is
    (n_energy, n_comp) on ascending energies; a grid that is not strictly
    ascending raises."""
    e = np.asarray(energy, dtype=np.float64)
    s = np.asarray(sigma_spectrum, dtype=np.float64)
    if s.shape[0] != e.size:
        raise ValueError("integrate_dc_spectrum: the spectrum and the energy grid differ in length")
    n_comp = s.shape[1] if s.ndim > 1 else 1
    if e.size == 0:
        return np.zeros(n_comp)
    if temp <= 0.0:
        return interpolate_spectrum_component(e, s, mu)
    de = np.diff(e)
    if not np.all(de > 0.0):
        raise RuntimeError("integrate_dc_spectrum: energies must be strictly ascending")
    kt = KB_EV * temp
    x = (e - mu) / kt
    softplus = np.maximum(x, 0.0) + np.log1p(np.exp(-np.abs(x)))
    F = e - kt * softplus
    f = fermi_dirac_energy(e, mu, temp)
    weight = f[:-1] - f[1:]
    moment = F[1:] - F[:-1] - f[1:] * de
    s0 = s[:-1].reshape(len(de), -1)
    s1 = s[1:].reshape(len(de), -1)
    out = (s0 * weight[:, None] + (s1 - s0) / de[:, None] * moment[:, None]).sum(axis=0)
    return out if s.ndim > 1 else float(out[0])


def fermi_window_nodes(n_moments: int, center: float, half_width: float, mu: float, temp: float):
    """Indices k of the ascending Chebyshev energy grid within
    max(20 kT, 3 pi half_width / n_moments) of mu: the nodes that -df/dE
    weighs (rsatb/calc/observables.hpp:812-828). Raises if none is."""
    window = max(20.0 * temp * KB_EV, 3.0 * math.pi * half_width / n_moments)
    k = np.arange(n_moments)
    theta = math.pi * ((n_moments - 1 - k) + 0.5) / n_moments
    energy = center + half_width * np.cos(theta)
    nodes = np.nonzero(np.abs(energy - mu) <= window)[0]
    if nodes.size == 0:
        raise RuntimeError("dc: mu lies outside the spectral window")
    return nodes
