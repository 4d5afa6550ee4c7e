"""FE workload generators: sizes of the named configurations and the closed
forms the matrices must satisfy (CPU)."""
import numpy as np
import pytest

from warpkernel_amd import fe


def test_q1_2d_sizes_and_nullspace():
    m = fe.q1_laplace_2d(128)
    assert (m.nrows, m.nnz) == (16641, 148225)
    A = m.to_scipy()
    assert abs(A @ np.ones(m.nrows)).max() < 1e-13 * abs(m.vals).max()
    assert abs(A - A.T).max() == 0


def test_q1_element_closed_form():
    K = fe.q1_element(2, 1)
    # bilinear square: 2/3 on the diagonal, -1/6 along edges, -1/3 across
    assert np.allclose(np.diag(K), 2 / 3)
    assert np.isclose(K[0, 1], -1 / 6) and np.isclose(K[0, 2], -1 / 6) and np.isclose(K[0, 3], -1 / 3)
    K3 = fe.q1_element(3, 1)
    assert np.allclose(np.diag(K3), 1 / 3)
    assert np.isclose(K3[0, 1], 0, atol=1e-15) and np.isclose(K3[0, 3], -1 / 12) and np.isclose(K3[0, 7], -1 / 12)


def test_hex_diffusion_sizes():
    m = fe.hex_laplace_3d(48)
    assert (m.nrows, m.nnz) == (117649, 3048625)
    assert abs(m.to_scipy() @ np.ones(m.nrows)).max() < 1e-12


def test_hex_elasticity_rigid_modes():
    n = 12
    m = fe.hex_elastic_3d(n)
    A = m.to_scipy()
    nodes = (n + 1) ** 3
    idx = np.arange(nodes)
    X, Y, Z = idx % (n + 1), (idx // (n + 1)) % (n + 1), idx // (n + 1) ** 2
    scale = abs(m.vals).max() * n
    for d in range(3):
        t = np.zeros(m.nrows)
        t[d::3] = 1
        assert abs(A @ t).max() < 1e-12 * scale
    for (a, b) in [(0, 1), (1, 2), (0, 2)]:
        P = [X, Y, Z]
        u = np.zeros(m.nrows)
        u[a::3] = -P[b]
        u[b::3] = P[a]
        assert abs(A @ u).max() < 1e-12 * scale
    # and it is not singular beyond the rigid modes: a random field has energy
    v = np.random.default_rng(0).uniform(-1, 1, m.nrows)
    assert v @ (A @ v) > 0


def test_hex_elasticity_64_census():
    m = fe.hex_elastic_3d(64)
    assert (m.nrows, m.nnz) == (823875, 64701513)
    u, c = np.unique(m.row_lengths(), return_counts=True)
    assert dict(zip(u.tolist(), c.tolist())) == {24: 24, 36: 2268, 54: 71442, 81: 750141}


def test_tet_poisson_structure():
    K = fe.tet_elements()
    for t in range(6):
        assert np.allclose(K[t].sum(axis=1), 0, atol=1e-14)
        assert np.allclose(K[t], K[t].T)
    m = fe.tet_poisson_3d(16, drop=0.0)
    lens = m.row_lengths()
    assert lens.max() == 15 and m.nrows == 17 ** 3
    A = m.to_scipy()
    assert abs(A @ np.ones(m.nrows)).max() < 1e-13
    md = fe.tet_poisson_3d(16, drop=0.05)
    assert md.nnz < m.nnz and len(np.unique(md.row_lengths())) > len(np.unique(lens))


def test_permutation_is_similarity():
    m = fe.tet_poisson_3d(8, drop=0.05)
    perm = np.random.default_rng(1).permutation(m.nrows).astype(np.int32)
    p = fe.permute(m, perm)
    A, B = m.to_scipy().toarray(), p.to_scipy().toarray()
    assert np.array_equal(B[np.ix_(perm, perm)], A)
    assert np.all(np.diff(p.colind[p.rowptr[0]:p.rowptr[1]]) > 0)


def test_batch_mix_is_seeded_and_spread():
    ms = fe.batch_mix(16, seed=77)
    ms2 = fe.batch_mix(16, seed=77)
    assert [m.nrows for m in ms] == [m.nrows for m in ms2]
    rows = np.array([m.nrows for m in ms])
    assert rows.min() >= 1000 and rows.max() <= 400000
    assert {m.name.split("_")[0] for m in ms} <= {"q1", "hex"}
