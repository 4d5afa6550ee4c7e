"""Adjoint through torch.autograd: dx = A^T dy and dvals against the oracle's
FP64 reference gradient and against central finite differences."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from warpkernel_amd import fe

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def reference_gradient(m, vals, x, b):
    """L = 0.5*||A(vals) x - b||^2: r = A x - b, dL/dx = A^T r, dL/dvals = r[row] x[col]."""
    r = O.csr_spmv(m.rowptr, m.colind, vals, x) - b
    cp, ri, perm = O.transpose(m.nrows, m.ncols, m.rowptr, m.colind)
    gx = O.csr_spmv(cp, ri, vals[perm], r)
    gv = O.dvals(m.rowptr, m.colind, r, x)
    return gx, gv


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("k", [1, 8])
@pytest.mark.parametrize("case", ["hex_laplace_3d_48", "tet_random_values"])
def test_adjoint_matches_reference_and_fd(case, k):
    from warpkernel_amd.autograd import WpkOperator
    say(f"[gpu] autograd {case} k={k}")
    if case == "hex_laplace_3d_48":
        m = fe.hex_laplace_3d(48)
        vals = m.vals
    else:  # non-symmetric values on a tet pattern: A^T != A
        m = fe.tet_poisson_3d(40, drop=0.05)
        vals = np.random.default_rng(6000).uniform(-1, 1, m.nnz)
    shape = (m.ncols,) if k == 1 else (m.ncols, k)
    x = fe.uniform_x(m.ncols, k, seed=2000 + k).reshape(shape)
    b = fe.uniform_x(m.nrows, k, seed=3000 + k).reshape((m.nrows,) + shape[1:])
    op = WpkOperator(m.rowptr, m.colind, m.vals)  # built with other values: forward re-syncs
    tv = torch.from_numpy(vals.copy()).cuda().requires_grad_(True)
    tx = torch.from_numpy(x.copy()).cuda().requires_grad_(True)
    tb = torch.from_numpy(b).cuda()
    y = op(tv, tx)
    loss = 0.5 * ((y - tb) ** 2).sum()
    loss.backward()
    gx_ref, gv_ref = reference_gradient(m, vals, x, b)
    gx, gv = tx.grad.cpu().numpy(), tv.grad.cpu().numpy()
    assert rel(gx, gx_ref) <= 1e-12 and rel(gv, gv_ref) <= 1e-12
    if k == 1:
        assert np.array_equal(gx, gx_ref) and np.array_equal(gv, gv_ref)
    # central differences on 256 coordinates each, FP64 on the GPU; L is
    # quadratic in vals and in x, so the difference quotient is exact up to rounding
    rng = np.random.default_rng(6000 + k)
    h = 1e-2

    def L(vv, xx):
        with torch.no_grad():
            return float(0.5 * ((op(vv, xx) - tb) ** 2).sum())

    with torch.no_grad():
        v0, x0 = tv.detach().clone(), tx.detach().clone()
    for j in rng.choice(m.nnz, 256, replace=False):
        vp, vm = v0.clone(), v0.clone()
        vp[j] += h
        vm[j] -= h
        fd = (L(vp, x0) - L(vm, x0)) / (2 * h)
        assert abs(fd - gv[j]) <= 1e-6 * max(1.0, abs(gv[j])), (j, fd, gv[j])
    flat = x0.view(-1)
    for j in rng.choice(flat.numel(), 256, replace=False):
        xp, xm = flat.clone(), flat.clone()
        xp[j] += h
        xm[j] -= h
        fd = (L(v0, xp.view(x0.shape)) - L(v0, xm.view(x0.shape))) / (2 * h)
        assert abs(fd - gx.reshape(-1)[j]) <= 1e-6 * max(1.0, abs(gx.reshape(-1)[j])), (j, fd)


def test_values_change_between_steps():
    """A training-style loop: new values every step go through the GPU-side
    process; results follow the values."""
    from warpkernel_amd.autograd import WpkOperator
    say("[gpu] autograd value updates")
    m = fe.hex_laplace_3d(16)
    op = WpkOperator(m.rowptr, m.colind, m.vals)
    x = torch.from_numpy(fe.uniform_x(m.ncols, seed=1)).cuda()
    tv = torch.from_numpy(m.vals.copy()).cuda().requires_grad_(True)
    opt = torch.optim.SGD([tv], lr=1e-3)
    for step in range(3):
        opt.zero_grad()
        y = op(tv, x)
        ref = O.csr_spmv(m.rowptr, m.colind, tv.detach().cpu().numpy(), x.cpu().numpy())
        assert np.array_equal(y.detach().cpu().numpy(), ref)
        (y ** 2).sum().backward()
        opt.step()


@pytest.mark.parametrize("variant,th", [(1, 0), (2, 3)])
@pytest.mark.parametrize("case", ["hex_laplace_3d_48", "tet_poisson_3d_96", "random_empty_rows"])
def test_backward_equals_adjoint_plus_dvals(case, variant, th):
    """wk_backward against run_adjoint + dvals, bit for bit, and dvals against
    the oracle."""
    from conftest import random_csr
    from warpkernel_amd import engine, structure, structure2
    say(f"[gpu] backward {case} variant={variant}")
    if case == "random_empty_rows":
        rp, cols, vals = random_csr(np.random.default_rng(5), 5000, 4000, 1)
        m = fe.CSR(case, 5000, 4000, rp, cols, vals)
    else:
        m = fe.CONFIGS[case]()
    s = (structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, ncols=m.ncols) if variant == 1 else
         structure2(64, 32).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, th, ncols=m.ncols))
    eng = engine(s, m.vals, m.colind, adjoint=True)
    rng = np.random.default_rng(6)
    for k in (1, 8):
        shp = lambda n: (n,) if k == 1 else (n, k)  # noqa: E731
        x = torch.from_numpy(rng.uniform(-1, 1, shp(m.ncols))).cuda()
        dy = torch.from_numpy(rng.uniform(-1, 1, shp(m.nrows))).cuda()
        dx1, dx2 = (torch.full(shp(m.ncols), float("nan"), dtype=torch.float64, device="cuda") for _ in range(2))
        dv1, dv2 = (torch.full((m.nnz,), float("nan"), dtype=torch.float64, device="cuda") for _ in range(2))
        eng.backward(dy, x, dx1, dv1)
        eng.run_adjoint(dy, dx2)
        eng.dvals(dy, x, dv2)
        torch.cuda.synchronize()
        assert torch.equal(dx1, dx2), k
        assert torch.equal(dv1, dv2), k
        assert np.array_equal(dv1.cpu().numpy(), O.dvals(m.rowptr, m.colind, dy.cpu().numpy(), x.cpu().numpy())), k
    eng.close()
