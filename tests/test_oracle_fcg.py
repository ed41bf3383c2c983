"""Pins for the oracle's flexible-CG pressure solver (oracle/pyoracle.py
mg_fcg_solve), which has no counterpart in the reference's code (the
reference names a PCG wrapper as the fix for boundary-limited multigrid,
docs/WEBGPU.md:462-465).  Pinned by mathematics:

* A (mg_apply_a) equals the 5-point matrix assembled here from the label
  rules (count on the diagonal, -1 per open fluid neighbour), and is symmetric;
* with the identity as preconditioner and CG from the start, FCG(1) is CG on
  an SPD system, so its iterates match an independent CG written here;
* its first cycle is exactly one defect-correction cycle;
* with the exact inverse as preconditioner it converges in one iteration;
* with the V-cycle it reaches the tolerance on open and closed boxes, agrees
  with the defect-correction solve, and needs fewer cycles than it.
"""
import numpy as np
import pytest

from oracle import pyoracle as O

SCALE = 0.02


def boundary_fixture(nx, ny):  # pressureMultigrid.test.ts:201-209
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[ny >> 2:ny >> 1, nx >> 2:(nx * 3) >> 3] = O.SOLID
    return lab.ravel(), O.lcg_fill(nx * ny, 12345)


def closed_fixture(nx, ny, seed=777):
    lab = np.zeros((ny, nx), np.uint8)
    lab[ny // 3:ny // 2, nx // 3:nx // 2] = O.SOLID
    lab = lab.ravel()
    div = O.lcg_fill(nx * ny, seed)
    div[lab != O.FLUID] = 0
    div[lab == O.FLUID] -= div[lab == O.FLUID].mean()  # compatible right-hand side
    return lab, div


def dense_a(nx, ny, lab):
    """The operator from the label rules: for a fluid cell, count = its
    non-solid in-domain neighbours; -1 couples it to each fluid neighbour;
    Air neighbours hold p = 0.  Cells that are not fluid, or fluid with no
    open neighbour, have an empty row."""
    L = lab.reshape(ny, nx)
    n = nx * ny
    A = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            if L[j, i] != O.FLUID:
                continue
            k = i + j * nx
            nbs = [(i + 1, j), (i - 1, j), (i, j + 1), (i, j - 1)]
            open_ = [(a, b) for a, b in nbs if 0 <= a < nx and 0 <= b < ny and L[b, a] != O.SOLID]
            if not open_:
                continue
            A[k, k] = len(open_)
            for a, b in open_:
                if L[b, a] == O.FLUID:
                    A[k, a + b * nx] = -1.0
    return A


def active(nx, ny, lab):
    return np.abs(dense_a(nx, ny, lab)).sum(axis=1) > 0


def test_apply_a_is_the_label_operator():
    nx, ny = 13, 9
    lab, _ = boundary_fixture(nx, ny)
    A = dense_a(nx, ny, lab)
    assert np.array_equal(A, A.T)
    x = O.lcg_fill(nx * ny, 99) * active(nx, ny, lab)
    assert np.allclose(O.mg_apply_a(nx, ny, x, lab), A @ x, rtol=0, atol=1e-13)


def test_identity_preconditioner_is_textbook_cg():
    nx, ny = 16, 12
    lab, div = boundary_fixture(nx, ny)
    act = active(nx, ny, lab)
    div = div * act
    A = dense_a(nx, ny, lab)
    b = -SCALE * div
    k = 12
    p_f, it, _ = O.mg_fcg_solve(nx, ny, np.zeros(nx * ny), div, lab, SCALE, k, 0,
                                precond=lambda r: r * act, dc_first=False)
    assert it == k
    # independent CG (Hestenes-Stiefel) on the same system
    x = np.zeros(nx * ny)
    r = b.copy()
    d = r.copy()
    for _ in range(k):
        q = A @ d
        a = (r @ r) / (d @ q)
        x += a * d
        rn = r - a * q
        d = rn + (rn @ rn) / (r @ r) * d
        r = rn
    assert np.linalg.norm(p_f - x) <= 1e-9 * np.linalg.norm(x)


def test_exact_preconditioner_converges_in_one_iteration():
    nx, ny = 12, 10
    lab, div = boundary_fixture(nx, ny)
    act = active(nx, ny, lab)
    div = div * act
    Ainv = np.linalg.pinv(dense_a(nx, ny, lab))
    p, it, hist = O.mg_fcg_solve(nx, ny, np.zeros(nx * ny), div, lab, SCALE, 5, 1e-10,
                                 precond=lambda r: Ainv @ r)
    assert it == 1 and hist[-1] < 1e-10


@pytest.mark.parametrize("closed", [False, True])
def test_vcycle_preconditioned_solve(closed):
    nx, ny = 96, 64
    lab, div = closed_fixture(nx, ny) if closed else boundary_fixture(nx, ny)
    p, it, hist = O.mg_fcg_solve(nx, ny, np.zeros(nx * ny), div, lab, SCALE, 60, 1e-6)
    assert hist[-1] <= 1e-6 and len(hist) == it + 1
    assert O.relative_residual(nx, ny, p, div, lab, SCALE) <= 1e-6
    pd, it_dc, _ = O.mg_solve(nx, ny, np.zeros(nx * ny), div, lab, SCALE, 1, 60, 1e-6)
    assert it < it_dc
    fl = lab == O.FLUID
    if closed:  # pure Neumann: p is defined up to a constant
        p = p - p[fl].mean()
        pd = pd - pd[fl].mean()
    assert np.linalg.norm((p - pd)[fl]) <= 1e-4 * np.linalg.norm(pd[fl])


def test_first_step_is_one_defect_correction_cycle():
    nx, ny = 48, 32
    lab, div = boundary_fixture(nx, ny)
    p0 = O.lcg_fill(nx * ny, 5) * (lab == O.FLUID)
    p1, it, _ = O.mg_fcg_solve(nx, ny, p0, div, lab, SCALE, 1, 0)
    assert it == 1
    exp = O.mg_dc_cycle(nx, ny, p0, -SCALE * div, lab)
    assert np.allclose(p1, exp, rtol=0, atol=1e-12 * np.abs(exp).max())


def test_zero_divergence_guard():
    nx, ny = 24, 16
    lab, _ = boundary_fixture(nx, ny)
    p, it, hist = O.mg_fcg_solve(nx, ny, np.ones(nx * ny), np.zeros(nx * ny), lab, SCALE, 10,
                                 1e-6)
    assert it == 0 and not p.any() and len(hist) == 0
