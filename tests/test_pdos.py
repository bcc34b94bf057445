"""Projected DOS moments on the CPU: the oracle's restatement of
kpm_pdos_moments (rsatb/calc/pdos.hpp:149-266) against the dense group-
resolved recurrence on a complete basis (where the estimator is exact), the
sum rule, mu_0 of +-1 samples, and the row groupings (pdos.hpp:56-127) against
the cases of rsatb/calc/test/test_pdos.cpp."""
import numpy as np
import pytest

from pdos_cases import K_HI, K_LO, N_ATOMS, NORB, as_bsr, chain
from rsatb_amd.pdos import pdos_groups_by_atom, pdos_groups_by_orbital_slot


def test_multi_species_grouping():
    """test_pdos.cpp:51-104: alternating C and H; types ascending by atomic
    number (H = 0 with 2 orbitals, C = 1 with 4): 6 groups, never merged."""
    n = 12
    atom_type = np.array([1 if i % 2 == 0 else 0 for i in range(n)])
    g = pdos_groups_by_orbital_slot(atom_type, [2, 4])
    assert g.n_groups == 6
    assert list(g.group_type) == [0, 0, 1, 1, 1, 1]
    assert list(g.group_slot) == [0, 1, 0, 1, 2, 3]
    row = 0
    for a in range(n):
        t = atom_type[a]
        for o in range((2, 4)[t]):
            gid = g.group_of_row[row]
            assert g.group_type[gid] == t and g.group_slot[gid] == o
            row += 1
    assert row == g.group_of_row.size


def test_single_species_groupings():
    """test_pdos.cpp:240-258: one species, by atom = p // norb, by slot = p % norb."""
    at = np.zeros(N_ATOMS, dtype=int)
    ga = pdos_groups_by_atom(at, [NORB])
    gs = pdos_groups_by_orbital_slot(at, [NORB])
    p = np.arange(N_ATOMS * NORB)
    assert ga.n_groups == N_ATOMS and np.array_equal(ga.group_of_row, p // NORB)
    assert gs.n_groups == NORB and np.array_equal(gs.group_of_row, p % NORB)
    assert list(gs.group_slot) == list(range(NORB)) and list(gs.group_type) == [0] * NORB
    # atom ids follow the caller's input index, not the storage order
    perm = np.random.default_rng(3).permutation(N_ATOMS)
    gp = pdos_groups_by_atom(at, [NORB], perm)
    assert np.array_equal(gp.group_of_row, np.repeat(perm, NORB))


@pytest.mark.parametrize("grouping", ["atom", "slot"])
def test_oracle_pdos_complete_basis_matches_dense(oracle, grouping):
    """test_pdos.cpp:262-298 (sandwich / Lowdin): with the identity as the
    sample block, N * mu^g equals sum_{i in g} T_m(Z^T H Z scaled)_ii."""
    H, S = chain()
    Z = as_bsr(oracle.isqrt_patches(S))
    N = N_ATOMS * NORB
    M = 40
    gor = np.arange(N) // NORB if grouping == "atom" else np.arange(N) % NORB
    G = N_ATOMS if grouping == "atom" else NORB
    op = oracle.Operator(1, H, 0.5 * (K_HI + K_LO), 0.5 * (K_HI - K_LO), Z=Z)
    total, group = op.pdos_moments(np.eye(N), M, gor, G)
    A = oracle.dense_scaled(1, H, 0.5 * (K_HI + K_LO), 0.5 * (K_HI - K_LO), Z).real
    ref = oracle.dense_group_moments(A, gor, G, M)
    assert np.max(np.abs(N * group.real - ref[:G])) < 1e-8
    assert np.max(np.abs(N * total.real - ref[G])) < 1e-8
    assert np.max(np.abs(group.sum(axis=0) - total)) < 1e-12
    # the plain recurrence gives the same moments (doubling is exact)
    t2, g2 = op.pdos_moments(np.eye(N), M, gor, G, doubling=False)
    assert np.max(np.abs(g2 - group)) < 1e-10
    # and the all-rows channel is the DOS moment loop's result
    assert np.max(np.abs(total - op.moments(np.eye(N), M))) < 1e-13


def test_oracle_pdos_orthogonal_complete_basis(oracle):
    """Orthogonal form on the chain's H: same check with the plain operator."""
    H, _ = chain()
    N = N_ATOMS * NORB
    M = 33
    gor = np.arange(N) // NORB
    op = oracle.Operator(0, H, 0.5, 11.5)
    total, group = op.pdos_moments(np.eye(N), M, gor, N_ATOMS)
    A = oracle.dense_scaled(0, H, 0.5, 11.5).real
    ref = oracle.dense_group_moments(A, gor, N_ATOMS, M)
    assert np.max(np.abs(N * group.real - ref[:N_ATOMS])) < 1e-10
    assert np.max(np.abs(N * total.real - ref[N_ATOMS])) < 1e-10


def test_oracle_pdos_rademacher_mu0(oracle):
    """test_pdos.cpp:334-349: (+-1)^2 = 1, so mu_0^g = |g| sample by sample."""
    H, S = chain()
    Z = as_bsr(oracle.isqrt_patches(S))
    N = N_ATOMS * NORB
    rng = np.random.default_rng(7)
    v = rng.choice([-1.0, 1.0], size=(N, 8))
    op = oracle.Operator(1, H, 0.0, 12.0, Z=Z)
    total, group = op.pdos_moments(v, 4, np.arange(N) // NORB, N_ATOMS)
    assert np.max(np.abs(group[:, 0].real - NORB)) < 1e-12
    assert abs(total[0].real - N) < 1e-12


def test_oracle_pdos_partial_grouping_and_errors(oracle):
    """Rows in no group (-1) leave the total unchanged and drop out of the
    groups; an id >= n_groups is an error (pdos.hpp:163-167)."""
    H, _ = chain()
    N = N_ATOMS * NORB
    op = oracle.Operator(0, H, 0.0, 12.0)
    psi = oracle.samples_rademacher(5, N, 4)
    gor = np.where(np.arange(N) % 5 == 0, -1, np.arange(N) % 2)
    total, group = op.pdos_moments(psi, 17, gor, 3)
    assert np.max(np.abs(total - op.moments(psi, 17))) < 1e-13
    assert np.all(group[2] == 0)
    full = np.arange(N) % 2
    _, gfull = op.pdos_moments(psi, 17, np.where(gor < 0, 2, full), 3)
    assert np.max(np.abs(gfull.sum(axis=0) - total)) < 1e-13
    with pytest.raises(RuntimeError, match="exceeds"):
        op.pdos_moments(psi, 4, np.full(N, 3), 3)
