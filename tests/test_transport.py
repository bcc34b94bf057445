"""Finite-temperature DC integral (rsatb_amd.transport) on the CPU: the
reference's exactness test (rsatb/calc/test/test_fftcheb_transport.cpp:747-766),
an independent quadrature, the T = 0 limit and the qe node window."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import kg_oracle as kg
from rsatb_amd import transport as tr


def _grid():
    return -11.3 + 1.13 * np.arange(21)


@pytest.mark.parametrize("offset", [0.0, 0.1, 0.37, 0.5, 0.81, 1.0])
def test_temperature_integral_exact_between_nodes(offset):
    """A spectrum linear in E over the whole grid integrates to its value at
    mu, a constant one to the constant, to 1e-9 relative, wherever mu falls
    between two nodes (test_fftcheb_transport.cpp:747-766)."""
    energy = _grid()
    a, b = 1.7e6, 2.0e5
    linear = np.repeat((a + b * energy)[:, None], 6, axis=1)
    const = np.full((21, 6), a)
    mu = energy[10] + offset * 1.13
    out_lin = tr.integrate_dc_spectrum(linear, energy, mu, 300.0)
    out_const = tr.integrate_dc_spectrum(const, energy, mu, 300.0)
    np.testing.assert_allclose(out_lin, a + b * mu, rtol=0, atol=1e-9 * a)
    np.testing.assert_allclose(out_const, a, rtol=0, atol=1e-9 * a)


def test_descending_energies_raise():
    energy = _grid()[::-1].copy()
    with pytest.raises(RuntimeError, match="strictly ascending"):
        tr.integrate_dc_spectrum(np.ones((21, 6)), energy, 0.0, 300.0)
    with pytest.raises(RuntimeError, match="strictly ascending"):
        tr.integrate_dc_spectrum(np.ones((3, 6)), np.array([0.0, 1.0, 1.0]), 0.5, 300.0)


@pytest.mark.parametrize("temp", [30.0, 300.0, 3000.0, 20000.0])
def test_temperature_integral_matches_quadrature(temp):
    """A rough spectrum whose piecewise-linear integral against -df/dE is
    checked by composite Simpson on a fine subdivision: a dropped f(e1) de
    term or a wrong softplus sign fails by far more than the tolerance."""
    rng = np.random.default_rng(7)
    energy = np.sort(rng.uniform(-4.0, 4.0, 40))
    spec = rng.normal(size=(40, 6)) * 1e5 + 1e6
    for mu in (-1.234, 0.0, 0.777, energy[17]):
        got = tr.integrate_dc_spectrum(spec, energy, mu, temp)
        ref = kg.dc_temperature_quadrature(spec, energy, mu, temp, per_segment=2000)
        np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-6)


def test_zero_temperature_interpolates():
    """At temp <= 0 the result is the spectrum interpolated linearly at mu,
    clamped to the end values (fftcheb_transport.hpp:950-1013); a vanishing
    but positive temperature approaches the same value."""
    energy = np.array([-1.0, 0.0, 0.5, 2.0])
    spec = np.array([[1.0, 2.0], [3.0, -1.0], [5.0, 0.0], [-2.0, 4.0]])
    np.testing.assert_allclose(tr.integrate_dc_spectrum(spec, energy, 0.25, 0.0), [4.0, -0.5])
    np.testing.assert_allclose(tr.integrate_dc_spectrum(spec, energy, 0.0, 0.0), [3.0, -1.0])
    np.testing.assert_allclose(tr.integrate_dc_spectrum(spec, energy, -5.0, 0.0), spec[0])
    np.testing.assert_allclose(tr.integrate_dc_spectrum(spec, energy, 9.0, -1.0), spec[-1])
    np.testing.assert_allclose(tr.integrate_dc_spectrum(spec, energy, 0.25, 1.0), [4.0, -0.5], atol=1e-6)
    assert np.all(tr.integrate_dc_spectrum(np.zeros((0, 6)), np.zeros(0), 0.0, 300.0) == 0.0)


def test_fermi_dirac_limits():
    e = np.array([-50.0, -0.1, 0.0, 0.1, 50.0])
    f = tr.fermi_dirac_energy(e, 0.0, 300.0)
    kt = tr.KB_EV * 300.0
    np.testing.assert_allclose(f, 1.0 / (1.0 + np.exp(np.clip(e / kt, -700, 700))), rtol=1e-14, atol=1e-300)
    assert f[2] == 0.5
    np.testing.assert_array_equal(tr.fermi_dirac_energy(e, 0.0, 0.0), [1, 1, 1, 0, 0])


def test_qe_window_nodes():
    """The nodes kept by the windowed route are those of the ascending
    Chebyshev grid within max(20 kT, 3 pi a / M) of mu
    (rsatb/calc/observables.hpp:812-828)."""
    M, c, a = 64, 0.3, 5.0
    energy = kg.energies(M, c, a)
    for mu, temp in ((0.0, 300.0), (2.0, 20000.0), (-4.9, 10.0)):
        nodes = tr.fermi_window_nodes(M, c, a, mu, temp)
        window = max(20 * tr.KB_EV * temp, 3 * math.pi * a / M)
        np.testing.assert_array_equal(nodes, np.nonzero(np.abs(energy - mu) <= window)[0])
        assert len(nodes) >= 1
    with pytest.raises(RuntimeError, match="outside the spectral window"):
        tr.fermi_window_nodes(M, c, a, 50.0, 300.0)


def test_twisted_bilayer_graphene_cell():
    """The (31, 32) commensurate cell: 5954 atoms per layer (2 (m^2 + m n + n^2)),
    a hexagonal superlattice of side a / (2 sin(phi / 2)) with phi = 1.05
    degrees (cos phi = 5953 / 5954), and both layers periodic in it: every
    atom has exactly three in-layer neighbours at 1.42 A, the bond
    directions of the top layer turned by phi against the bottom's."""
    from rsatb_amd.lattice import neighbor_edges, twisted_bilayer_graphene

    st = twisted_bilayer_graphene(31, 32)
    assert st.n_atoms == 2 * 5954
    assert np.count_nonzero(st.atom_type == 0) == 5954
    phi = math.acos(5953 / 5954)
    assert abs(math.degrees(phi) - 1.050) < 1e-3
    a = math.sqrt(3) * 1.42
    side = a / (2 * math.sin(phi / 2))
    np.testing.assert_allclose(np.linalg.norm(st.cell[:2], axis=1), side, rtol=1e-12)
    assert abs(st.cell[0] @ st.cell[1] - 0.5 * side * side) < 1e-8 * side * side
    e = neighbor_edges(st, 1.6)
    same = st.atom_type[e.src] == st.atom_type[e.dst]
    assert np.all(same)
    np.testing.assert_allclose(e.dist, 1.42, atol=1e-9)
    np.testing.assert_array_equal(np.bincount(e.src, minlength=st.n_atoms), 3)
    ang = np.arctan2(e.vec[:, 1], e.vec[:, 0]) % (math.pi / 3)  # bonds of both sublattices
    bottom = np.unique(np.round(ang[st.atom_type[e.src] == 0], 9))
    top = np.unique(np.round(ang[st.atom_type[e.src] == 1], 9))
    assert bottom.size == 1 and top.size == 1
    d = (top[0] - bottom[0]) % (math.pi / 3)
    assert abs(min(d, math.pi / 3 - d) - phi) < 1e-9
