"""Seeded finite-element SpMV workloads (host CSR, int32 indices, FP64 values).

Matrices are assembled by libwk_fegen.so straight into CSR with sorted columns;
the sparsity pattern is the element connectivity, exact zeros kept.

Workloads (seeds: coefficients 1000+c, x 2000+c, b 3000+c, element drop
4000+c, node permutation 5000+c; numpy default_rng):
  q1_laplace_2d(128)          16,641 rows, 148,225 nnz        (configuration 1)
  hex_elastic_3d(64)          823,875 rows, 64,701,513 nnz    (configuration 2)
  tet_poisson_3d(96, ...)     912,673 rows, 5% elements dropped (configuration 3)
  hex_elastic_3d(32)          107,811 rows, 8,214,057 nnz     (configuration 4)
  batch_mix(64)               64 mixed 2D-Q1 / 3D-hex matrices (configuration 4b)
  hex_laplace_3d(48)          117,649 rows, 3,048,625 nnz     (configuration 5)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_PKG, "libwk_fegen.so")
_lib = None


def _gen():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"FE generator not built: {_LIB_PATH}")
        L = C.CDLL(_LIB_PATH)
        P, I = C.c_void_p, C.c_int
        L.fe_q1_rowptr.restype = C.c_int64
        L.fe_q1_rowptr.argtypes = [I, I, I, I, I, P]
        L.fe_q1_fill.argtypes = [I, I, I, I, I, C.c_double, P, P, P, P]
        L.fe_q1_element.argtypes = [I, I, C.c_double, P]
        L.fe_tet_rowptr.restype = C.c_int64
        L.fe_tet_rowptr.argtypes = [I, I, I, P, P]
        L.fe_tet_fill.argtypes = [I, I, I, P, P, P, P, P]
        L.fe_tet_element.argtypes = [P]
        L.fe_permute_rowptr.argtypes = [C.c_int32, P, P, P]
        L.fe_permute_fill.argtypes = [C.c_int32, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class CSR:
    name: str
    nrows: int
    ncols: int
    rowptr: np.ndarray
    colind: np.ndarray
    vals: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.rowptr)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.colind, self.rowptr), shape=(self.nrows, self.ncols))


def _q1(name, dim, ex, ey, ez, ndof, coef, nu=0.3) -> CSR:
    L = _gen()
    nodes = (ex + 1) * (ey + 1) * ((ez + 1) if dim == 3 else 1)
    n = nodes * ndof
    rowptr = np.empty(n + 1, np.int32)
    nnz = L.fe_q1_rowptr(dim, ex, ey, ez, ndof, _p(rowptr))
    colind = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    L.fe_q1_fill(dim, ex, ey, ez, ndof, nu, _p(coef), _p(rowptr), _p(colind), _p(vals))
    return CSR(name, n, n, rowptr, colind, vals)


def q1_laplace_2d(n: int = 128, seed: int = 1001) -> CSR:
    """Bilinear Q1 Laplacian on an n x n element grid, kappa ~ lognormal(0, 0.5)."""
    kappa = np.random.default_rng(seed).lognormal(0.0, 0.5, n * n)
    m = _q1(f"q1_laplace_2d_{n}", 2, n, n, 0, 1, kappa)
    m.meta.update(kappa=kappa, seed=seed)
    return m


def hex_laplace_3d(n: int = 48, seed: int = 1005) -> CSR:
    """Trilinear Q1 scalar diffusion on an n^3 element grid (27-point)."""
    kappa = np.random.default_rng(seed).lognormal(0.0, 0.5, n ** 3)
    m = _q1(f"hex_laplace_3d_{n}", 3, n, n, n, 1, kappa)
    m.meta.update(kappa=kappa, seed=seed)
    return m


def hex_elastic_3d(n: int = 64, seed: int = 1002, nu: float = 0.3) -> CSR:
    """Trilinear Q1 isotropic linear elasticity on an n^3 grid, E ~ lognormal(0, 0.5),
    3 dof per node interleaved (row = 3*node + d)."""
    E = np.random.default_rng(seed).lognormal(0.0, 0.5, n ** 3)
    m = _q1(f"hex_elastic_3d_{n}", 3, n, n, n, 3, E, nu)
    m.meta.update(E=E, nu=nu, seed=seed)
    return m


def tet_poisson_3d(n: int = 96, drop: float = 0.05, seed: int = 1003, drop_seed: int = 4003,
                   perm_seed: int | None = None) -> CSR:
    """P1 Laplacian on the Kuhn subdivision of an n^3 cube grid (6 tets/cube),
    kappa ~ lognormal(0, 0.5) per tet, a fraction `drop` of the tets removed at
    random, optionally under a random symmetric node permutation."""
    L = _gen()
    ne = 6 * n ** 3
    keep = (np.random.default_rng(drop_seed).random(ne) >= drop).astype(np.uint8)
    kappa = np.random.default_rng(seed).lognormal(0.0, 0.5, ne)
    nrows = (n + 1) ** 3
    rowptr = np.empty(nrows + 1, np.int32)
    nnz = L.fe_tet_rowptr(n, n, n, _p(keep), _p(rowptr))
    colind = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    L.fe_tet_fill(n, n, n, _p(keep), _p(kappa), _p(rowptr), _p(colind), _p(vals))
    m = CSR(f"tet_poisson_3d_{n}", nrows, nrows, rowptr, colind, vals,
            meta=dict(keep=keep, kappa=kappa, seed=seed, drop_seed=drop_seed))
    if perm_seed is not None:
        perm = np.random.default_rng(perm_seed).permutation(nrows).astype(np.int32)
        m = permute(m, perm)
        m.name += "_perm"
        m.meta["perm"] = perm
    return m


def permute(m: CSR, perm: np.ndarray) -> CSR:
    """B = P A P^T with B[perm[i], perm[j]] = A[i, j]."""
    L = _gen()
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    rp = np.empty(m.nrows + 1, np.int32)
    L.fe_permute_rowptr(m.nrows, _p(m.rowptr), _p(perm), _p(rp))
    ci = np.empty(m.nnz, np.int32)
    va = np.empty(m.nnz, np.float64)
    L.fe_permute_fill(m.nrows, _p(m.rowptr), _p(m.colind), _p(m.vals), _p(perm), _p(rp), _p(ci), _p(va))
    return CSR(m.name, m.nrows, m.ncols, rp, ci, va, dict(m.meta))


def q1_element(dim: int, ndof: int, nu: float = 0.3) -> np.ndarray:
    n = (1 << dim) * ndof
    out = np.empty((n, n), np.float64)
    _gen().fe_q1_element(dim, ndof, nu, _p(out))
    return out


def tet_elements() -> np.ndarray:
    out = np.empty((6, 4, 4), np.float64)
    _gen().fe_tet_element(_p(out))
    return out


def batch_mix(count: int = 64, lo: int = 2000, hi: int = 250000, seed: int = 1004) -> list[CSR]:
    """`count` FE matrices with log-uniform target sizes in [lo, hi] rows, each
    a 2D Q1 Laplacian or a 3D hex elasticity matrix (coin flip), seeded."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        target = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if rng.random() < 0.5:
            n = max(2, int(round(np.sqrt(target))) - 1)
            out.append(q1_laplace_2d(n, seed=seed * 100 + i))
        else:
            n = max(1, int(round((target / 3.0) ** (1.0 / 3.0))) - 1)
            out.append(hex_elastic_3d(n, seed=seed * 100 + i))
    return out


def uniform_x(n: int, k: int = 1, seed: int = 2000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, n if k == 1 else (n, k))


# configuration name -> builder (host CSR)
CONFIGS = {
    "q1_laplace_2d_128": lambda: q1_laplace_2d(128),
    "hex_elastic_3d_64": lambda: hex_elastic_3d(64),
    "tet_poisson_3d_96": lambda: tet_poisson_3d(96),
    "tet_poisson_3d_96_perm": lambda: tet_poisson_3d(96, perm_seed=5003),
    "hex_elastic_3d_32": lambda: hex_elastic_3d(32),
    "hex_laplace_3d_48": lambda: hex_laplace_3d(48),
}
