"""Slater-Koster tables (host) and the SK block oracle (CPU).

Known answers are those of rsatb/model/sktb/test/test_sktb.cpp (line numbers
per test); closed forms pin what those leave open: the direction-cosine p-p
rule, the quadrupole d-d and p-d outer products, orthogonality of the shell
rotations, Hermiticity of assembled periodic matrices, and the p_z rule of
the host SkPzModel.
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from oracle import sk_oracle as so  # noqa: E402
from rsatb_amd.sktb import Radial, SkParameters, SkTable, parse_channel_name  # noqa: E402
from sk_cases import periodic_edges, silicon_cell, sp3_silicon_table  # noqa: E402

TOL = 1e-12


def constant_spline(v, cutoff=2.5):
    return Radial.spline([0.0, cutoff], [v, v], cutoff)


def two_atom_block(table, which, types, vec, directed=False):
    """Blocks (0,1) and (1,0) of two atoms at 0 and vec (non-periodic)."""
    pos = np.array([[0.0, 0.0, 0.0], vec], dtype=np.float64)
    ei = np.array([[0], [1]]) if directed else np.array([[0, 1], [1, 0]])
    sh = np.zeros((ei.shape[1], 3), dtype=np.int64)
    rp, ci, blk = so.assemble(table, which, np.array(types), pos, np.eye(3) * 10.0, ei, sh)
    out = {}
    for r in range(2):
        for s in range(rp[r], rp[r + 1]):
            out[(r, int(ci[s]))] = blk[s]
    return out


def carbon_p(sigma, pi, cutoff=2.5, overlap=False):
    h = SkParameters().set_onsite("C", "p", 0.0)
    h.set_pair_channel("C", "C", "pp", "sigma", constant_spline(sigma, cutoff))
    h.set_pair_channel("C", "C", "pp", "pi", constant_spline(pi, cutoff))
    s = None
    if overlap:
        s = SkParameters()
        s.set_pair_channel("C", "C", "pp", "sigma", constant_spline(sigma, cutoff))
        s.set_pair_channel("C", "C", "pp", "pi", constant_spline(pi, cutoff))
    return SkTable({"C": ["p"]}, h, s)


def p_sigma(n):
    return np.array([n[1], n[2], n[0]])


def d_sigma(n):
    x, y, z = n
    r3 = math.sqrt(3.0)
    return np.array([r3 * x * y, r3 * y * z, 0.5 * (3 * z * z - 1), r3 * z * x, 0.5 * r3 * (x * x - y * y)])


# ----------------------------------------------------------------- host table


def test_basis_expansion_and_atomic_preparation():
    """test_sktb.cpp:309-340."""
    h = SkParameters()
    h.onsite = {"C": {"pz": 0.0}, "Si": {"s": -8.0, "p": 0.0, "d": 6.0}}
    h.set_pair_channel("C", "Si", "sp", "sigma", constant_spline(-2.0, 3.7))
    t = SkTable({"C": ["pz"], "Si": ["s", "p", "d"]}, h)
    assert [o[0] for o in t.atom_types[0].orbitals] == ["pz"]
    assert [o[0] for o in t.atom_types[1].orbitals] == ["s", "py", "pz", "px", "dxy", "dyz", "dz2", "dzx",
                                                        "dx2-y2"]
    assert [a.symbol for a in t.atom_types] == ["C", "Si"]
    assert t.type_norb == [1, 9]
    assert t.cutoff_radii() == [1.85, 1.85]


def test_bond_cutoff_to_atomic_radius():
    """test_sktb.cpp:342-357."""
    h = SkParameters().set_onsite("C", "pz", 0.0)
    h.set_pair_channel("C", "C", "pp", "pi", constant_spline(-2.7, 8.26))
    h.set_pair_channel("C", "C", "pp", "sigma", constant_spline(0.3675, 8.26))
    assert SkTable({"C": ["pz"]}, h).cutoff_radii() == [4.13]


def test_duplicate_basis_labels_rejected():
    """test_sktb.cpp:359-367."""
    h = SkParameters().set_onsite("C", "p", 0.0)
    with pytest.raises(ValueError, match="duplicate"):
        SkTable({"C": ["p", "pz"]}, h)


def test_channel_names():
    assert parse_channel_name("pd_pi") == (1, 2, 1)
    assert parse_channel_name("dd_delta") == (2, 2, 2)
    for bad in ("ps_sigma", "sp_pi", "pp_delta", "xx_sigma", "pp"):
        with pytest.raises(ValueError):
            parse_channel_name(bad)


def test_table_kind_rejected():
    """test_sktb.cpp:392-402."""
    r = constant_spline(1.0, 2.0)
    r.kind = "table"
    with pytest.raises(ValueError, match="use 'spline'"):
        r.validate("legacy")


def test_uniform_block_arrays_and_rejection():
    t = carbon_p(5.0, 1.0)
    a = t.arrays("H")
    assert a["norb"] == 3 and a["n_types"] == 1 and a["n_channels"] == 2
    assert list(a["orb_l"]) == [1, 1, 1] and list(a["orb_m"]) == [0, 1, 2]
    h = SkParameters()
    h.onsite = {"C": {"pz": 0.0}, "Si": {"s": 0.0}}
    h.set_pair_channel("C", "Si", "sp", "sigma", constant_spline(1.0))
    SkTable({"C": ["pz"], "Si": ["s"]}, h).arrays("H")  # both 1 orbital: fine
    h2 = SkParameters()
    h2.onsite = {"C": {"pz": 0.0}, "Si": {"s": 0.0, "p": 0.0}}
    with pytest.raises(ValueError, match="uniform"):
        SkTable({"C": ["pz"], "Si": ["s", "p"]}, h2).arrays("H")


# ------------------------------------------------------------------- radials


def test_radial_functions():
    """test_sktb.cpp:369-390."""
    c = constant_spline(2.0, 3.0)
    assert so.radial_eval(c, 0.0) == 2.0
    assert so.radial_eval(c, 1.5) == 2.0
    assert so.radial_eval(c, 3.5) == 0.0
    sp = Radial.spline([0.0, 1.0, 2.0], [0.0, 1.0, 0.0], 2.0)
    assert so.radial_eval(sp, 0.0) == 0.0
    assert abs(so.radial_eval(sp, 1.0) - 1.0) < TOL
    assert so.radial_eval(sp, 2.5) == 0.0
    ep = Radial.exp_poly([4.0], 4.0, 0.0, 2.0)
    assert so.radial_eval(ep, 1.0) == 4.0
    assert abs(so.radial_eval(ep, 3.0) - 2.0) < TOL
    assert so.radial_eval(ep, 4.1) == 0.0


def test_spline_exact_for_lines_and_natural():
    """The natural cubic spline is exact for linear data, interpolates its
    knots, has y2 = 0 at both ends, and its y2 solve the continuity system
    h_{i-1} y2_{i-1} + 2 (h_{i-1} + h_i) y2_i + h_i y2_{i+1} = 6 (slope jump)."""
    r = np.array([0.5, 1.0, 1.7, 2.2, 3.0])
    lin = Radial.spline(r, 2.0 - 0.5 * r, 3.0)
    for d in (0.6, 1.3, 2.0, 2.9):
        assert abs(so.radial_eval(lin, d) - (2.0 - 0.5 * d)) < 1e-13
    vals = np.sin(r)
    s = Radial.spline(r, vals, 3.0)
    assert s.y2[0] == 0.0 and s.y2[-1] == 0.0
    for x, y in zip(r, vals):
        assert abs(so.radial_eval(s, x) - y) < 1e-13
    # the second derivatives solve the natural-spline continuity system
    h = np.diff(r)
    for i in range(1, r.size - 1):
        lhs = h[i - 1] * s.y2[i - 1] + 2 * (h[i - 1] + h[i]) * s.y2[i] + h[i] * s.y2[i + 1]
        rhs = 6 * ((vals[i + 1] - vals[i]) / h[i] - (vals[i] - vals[i - 1]) / h[i - 1])
        assert abs(lhs - rhs) < 1e-12


def test_exp_poly_closed_form():
    ep = Radial.exp_poly([1.0, -0.5, 0.25], 5.0, 0.7, 3.0)
    for d in (0.3, 2.9, 3.5, 4.9):
        w = 1.0 if d <= 3.0 else 0.5 * (1 + math.cos(math.pi * (d - 3.0) / 2.0))
        ref = w * math.exp(-0.7 * d) * (1.0 - 0.5 * d + 0.25 * d * d)
        assert abs(so.radial_eval(ep, d) - ref) < 1e-14


# ---------------------------------------------------------------- rotations


def test_rotations_are_orthogonal():
    rng = np.random.default_rng(5)
    for _ in range(20):
        f = so.local_frame(rng.normal(size=3))
        assert np.allclose(f.T @ f, np.eye(3), atol=1e-14)
        for l in (1, 2):
            R = so.shell_rotation(l, f)
            assert np.allclose(R @ R.T, np.eye(2 * l + 1), atol=1e-13)


def test_aligned_pd_block_magnetic_order():
    """test_sktb.cpp:680-697."""
    blk = so.aligned_shell_block(1, 2, 0, 6.0) + so.aligned_shell_block(1, 2, 1, 4.0)
    exp = np.zeros(15)
    exp[1], exp[7], exp[13] = 4.0, 6.0, 4.0
    assert np.array_equal(blk.ravel(), exp)


# ------------------------------------------------------------- pair blocks


def test_pp_along_bond():
    """test_sktb.cpp:404-419."""
    blk = two_atom_block(carbon_p(5.0, 1.0), "H", [0, 0], [1.0, 0.0, 0.0])
    assert np.allclose(blk[(0, 1)], np.diag([1.0, 1.0, 5.0]), atol=TOL)


def test_pp_axis_directions():
    """test_sktb.cpp:421-444."""
    t = carbon_p(5.0, 1.0)
    for vec, diag in (([1, 0, 0], [1, 1, 5]), ([0, 1, 0], [5, 1, 1]), ([0, 0, 1], [1, 5, 1])):
        blk = two_atom_block(t, "H", [0, 0], np.array(vec, float))
        assert np.allclose(blk[(0, 1)], np.diag(diag), atol=TOL)


def test_pp_skew_direction_cosines():
    """test_sktb.cpp:446-465: pi delta + (sigma - pi) p p^T."""
    v = np.array([1.0, 2.0, 3.0])
    blk = two_atom_block(carbon_p(5.0, 1.0, 10.0), "H", [0, 0], v)
    p = p_sigma(v / np.linalg.norm(v))
    assert np.allclose(blk[(0, 1)], np.eye(3) + 4.0 * np.outer(p, p), atol=TOL)


def _si_d(channels, cutoff=2.5):
    h = SkParameters().set_onsite("Si", "d", 0.0)
    for bond, v in channels.items():
        h.set_pair_channel("Si", "Si", "dd", bond, constant_spline(v, cutoff))
    return SkTable({"Si": ["d"]}, h)


def test_dd_along_z():
    """test_sktb.cpp:467-493."""
    blk = two_atom_block(_si_d({"sigma": 6.0, "pi": 4.0, "delta": 2.0}), "H", [0, 0], [0, 0, 1.5])
    assert np.allclose(blk[(0, 1)], np.diag([2.0, 4.0, 6.0, 4.0, 2.0]), atol=TOL)


def test_dd_sigma_skew_quadrupole():
    """test_sktb.cpp:495-519."""
    v = np.array([2.0, -1.0, 3.0])
    blk = two_atom_block(_si_d({"sigma": 6.0}, 10.0), "H", [0, 0], v)
    d = d_sigma(v / np.linalg.norm(v))
    assert np.allclose(blk[(0, 1)], 6.0 * np.outer(d, d), atol=TOL)


def test_sp_heteronuclear_transpose_consistent():
    """test_sktb.cpp:521-551."""
    h = SkParameters()
    h.onsite = {"H": {"s": 0.0}, "C": {"pz": 0.0}}
    h.set_pair_channel("H", "C", "sp", "sigma", constant_spline(3.0))
    t = SkTable({"H": ["s"], "C": ["pz"]}, h)
    # types are ordered by atomic number: H = 0, C = 1
    blk = two_atom_block(t, "H", [0, 1], [0, 0, 1.0])
    assert abs(blk[(0, 1)][0, 0] - 3.0) < TOL and abs(blk[(1, 0)][0, 0] - 3.0) < TOL


def test_sd_sigma_on_dz2():
    """test_sktb.cpp:553-584 (single s row of an s-d block)."""
    h = SkParameters()
    h.onsite = {"H": {"s": 0.0}, "Si": {"d": 0.0}}
    h.set_pair_channel("H", "Si", "sd", "sigma", constant_spline(7.0))
    t = SkTable({"H": ["s"], "Si": ["d"]}, h)
    blk = so.pair_block(t, "H", 0, 1, np.array([0, 0, 1.0]))
    assert np.allclose(blk, [[0, 0, 7.0, 0, 0]], atol=TOL)


def _c_si_pd(channels, cutoff=2.5):
    h = SkParameters()
    h.onsite = {"C": {"p": 0.0}, "Si": {"d": 0.0}}
    for bond, v in channels.items():
        h.set_pair_channel("C", "Si", "pd", bond, constant_spline(v, cutoff))
    return SkTable({"C": ["p"], "Si": ["d"]}, h)


def test_pd_aligned_z():
    """test_sktb.cpp:586-622."""
    blk = so.pair_block(_c_si_pd({"sigma": 6.0, "pi": 4.0}), "H", 0, 1, np.array([0, 0, 1.0]))
    exp = np.zeros(15)
    exp[1], exp[7], exp[13] = 4.0, 6.0, 4.0
    assert np.allclose(blk.ravel(), exp, atol=TOL)


def test_pd_sigma_skew_and_reverse_lookup():
    """test_sktb.cpp:624-678."""
    t = _c_si_pd({"sigma": 6.0}, 10.0)
    v = np.array([1.0, -2.0, 2.5])
    n = v / np.linalg.norm(v)
    direct = 6.0 * np.outer(p_sigma(n), d_sigma(n))
    assert np.allclose(so.pair_block(t, "H", 0, 1, v), direct, atol=TOL)
    assert np.allclose(so.pair_block(t, "H", 1, 0, -v), direct.T, atol=TOL)


def test_explicit_pz_selects_m0():
    """test_sktb.cpp:699-736."""
    h = SkParameters().set_onsite("C", "pz", 0.0)
    h.set_pair_channel("C", "C", "pp", "sigma", constant_spline(5.0))
    h.set_pair_channel("C", "C", "pp", "pi", constant_spline(1.0))
    t = SkTable({"C": ["pz"]}, h)
    assert abs(so.pair_block(t, "H", 0, 0, np.array([0, 0, 1.0]))[0, 0] - 5.0) < TOL
    assert abs(so.pair_block(t, "H", 0, 0, np.array([1.0, 0, 0]))[0, 0] - 1.0) < TOL


def test_overlap_separate_channels():
    """test_sktb.cpp:738-766."""
    h = SkParameters().set_onsite("C", "pz", 0.0)
    h.set_pair_channel("C", "C", "pp", "pi", constant_spline(0.20, 3.0))
    s = SkParameters().set_pair_channel("C", "C", "pp", "pi", constant_spline(0.05, 3.0))
    t = SkTable({"C": ["pz"]}, h, s)
    v = np.array([1.0, 0, 0])
    assert abs(so.pair_block(t, "H", 0, 0, v)[0, 0] - 0.20) < TOL
    assert abs(so.pair_block(t, "S", 0, 0, v)[0, 0] - 0.05) < TOL
    assert np.array_equal(t.atom_types[0].s_onsite, [1.0])


def test_beyond_cutoff_and_zero_blocks_dropped():
    t = carbon_p(5.0, 1.0, 2.0)
    assert so.pair_block(t, "H", 0, 0, np.array([2.5, 0, 0])) is None
    # a pz-only pi bond along z has an exactly zero block: dropped from the pattern
    h = SkParameters().set_onsite("C", "pz", 0.0)
    h.set_pair_channel("C", "C", "pp", "pi", constant_spline(1.0))
    tz = SkTable({"C": ["pz"]}, h)
    blk = two_atom_block(tz, "H", [0, 0], [0, 0, 1.0])
    assert set(blk) == {(0, 0), (1, 1)}


# ------------------------------------------------------------- assembly


def test_periodic_assembly_hermitian_and_k_phase():
    t = sp3_silicon_table()
    pos, cell = silicon_cell()
    ei, sh = periodic_edges(pos, cell, 3.2)
    types = np.zeros(len(pos), dtype=np.int32)
    for which in ("H", "S"):
        rp, ci, blk = so.assemble(t, which, types, pos, cell, ei, sh)
        n, b = len(pos), 9
        dense = np.zeros((n * b, n * b))
        for r in range(n):
            for s in range(rp[r], rp[r + 1]):
                dense[r * b:(r + 1) * b, ci[s] * b:(ci[s] + 1) * b] = blk[s]
        assert np.allclose(dense, dense.T, atol=1e-13)
        k = (0.13, -0.21, 0.37)
        rpk, cik, blkk = so.assemble(t, which, types, pos, cell, ei, sh, k)
        densek = np.zeros((n * b, n * b), dtype=complex)
        for r in range(n):
            for s in range(rpk[r], rpk[r + 1]):
                densek[r * b:(r + 1) * b, cik[s] * b:(cik[s] + 1) * b] = blkk[s]
        assert np.allclose(densek, densek.conj().T, atol=1e-13)


def test_pz_table_matches_host_pz_model():
    """A pz-only table reproduces SkPzModel's n^2 V_sigma + (1 - n^2) V_pi."""
    from rsatb_amd.lattice import graphene_rect_supercell, neighbor_edges
    from rsatb_amd.models import SkPzModel, assemble as host_assemble

    st = graphene_rect_supercell(3, 2)
    st.positions = st.positions.copy()
    st.positions[:, 2] += 0.3 * np.sin(np.arange(st.n_atoms))  # corrugate so n_z != 0
    e = neighbor_edges(st, 1.6)
    host = host_assemble(st, e, SkPzModel(onsite=0.25, shells=[(1.6, -2.7, 0.48)]), "H")
    h = SkParameters().set_onsite("C", "pz", 0.25)
    h.set_pair_channel("C", "C", "pp", "pi", constant_spline(-2.7, 1.6))
    h.set_pair_channel("C", "C", "pp", "sigma", constant_spline(0.48, 1.6))
    t = SkTable({"C": ["pz"]}, h)
    rp, ci, blk = so.assemble(t, "H", np.zeros(st.n_atoms, np.int32), st.positions, st.cell,
                              np.stack([e.src, e.dst]), e.shift)
    assert np.array_equal(rp, host.row_ptr) and np.array_equal(ci, host.col_ind)
    assert np.allclose(blk.ravel(), host.values.ravel(), atol=1e-13)
