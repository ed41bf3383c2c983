"""How much any coarse-grid correction of the solve's start can save, by the
FP64 oracle alone (no GPU).

The learned warm start predicts a coarse correction: the block means of the
error on F x F blocks, interpolated bilinearly onto the fine fluid cells.
Here that correction is given *exactly* (from the true solution p*), on top
of the start extrapolated linearly in time from the last two solutions, and
the V-cycles of flexible CG around the defect-correction V-cycle to 1e-6 are
counted.  On the plume scene of the 8192^2 headline (resolved at 256^2) the
exact coarse correction saves less than a cycle per solve over the
extrapolation, so no network of that kind can double the solve rate there;
the extrapolation itself saves cycles over the previous pressure.
"""
import numpy as np

from flowlab_amd import scenes as S
from oracle import pyoracle as O


def coarse_part(e, label, nx, ny, F):
    """Block means of e over the fluid cells of F x F blocks, interpolated
    bilinearly between block centres (clamped at the edges) onto the fine
    fluid cells: the form of the learned start's correction."""
    E = e.reshape(ny, nx)
    fl = (label.reshape(ny, nx) == O.FLUID).astype(float)
    s = (E * fl).reshape(ny // F, F, nx // F, F).sum((1, 3))
    c = fl.reshape(ny // F, F, nx // F, F).sum((1, 3))
    m = np.where(c > 0, s / np.maximum(c, 1), 0.0)
    cy, cx = m.shape
    y = np.clip((np.arange(ny) + 0.5) / F - 0.5, 0, cy - 1)
    x = np.clip((np.arange(nx) + 0.5) / F - 0.5, 0, cx - 1)
    y0 = np.minimum(y.astype(int), cy - 1)
    x0 = np.minimum(x.astype(int), cx - 1)
    y1, x1 = np.minimum(y0 + 1, cy - 1), np.minimum(x0 + 1, cx - 1)
    fy, fx = (y - y0)[:, None], (x - x0)[None, :]
    top = m[y0][:, x0] + fx * (m[y0][:, x1] - m[y0][:, x0])
    bot = m[y1][:, x0] + fx * (m[y1][:, x1] - m[y1][:, x0])
    return ((top + fy * (bot - top)) * fl).ravel()


def test_coarse_interpolation_reproduces_linear_fields():
    """The block mean of a linear field is its value at the block centre,
    and bilinear interpolation between centres is exact for it: every cell
    between the first and last block centres gets the field back."""
    nx, ny, F = 32, 16, 4
    label = np.full(nx * ny, O.FLUID, np.uint8)
    y, x = np.mgrid[0:ny, 0:nx] + 0.5
    e = (0.3 * x - 1.7 * y + 2.0).ravel()
    q = coarse_part(e, label, nx, ny, F).reshape(ny, nx)
    inner = (slice(F // 2, ny - F // 2), slice(F // 2, nx - F // 2))
    assert np.allclose(q[inner], e.reshape(ny, nx)[inner], rtol=0, atol=1e-12)
    # outside the outer block centres the interpolant is clamped to them
    assert np.allclose(q[:, 0], q[:, 1]) and np.allclose(q[0], q[1])
    # solids get no correction
    label[5] = O.SOLID
    assert coarse_part(np.ones(nx * ny), label, nx, ny, F)[5] == 0.0


def test_exact_coarse_correction_saves_under_a_cycle_on_the_plume():
    n, steps, first = 256, 16, 8
    g, scene = S.large_plume(n, cell_scale=n / 8192)
    sim = O.OracleSim(n, n)
    sim.reset(g, scene)
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=100, tol=1e-10,
                        cfl_target=0.5, dt_max=0.5 * g.h, residual_every=0, buoyancy=1.0,
                        patch=(sim.patch.i0, sim.patch.j0, sim.patch.nx, sim.patch.ny))
    ps, divs, scales, dts = [], [], [], []
    for _ in range(steps):
        sim.step(prm, sim.patch.data)
        ps.append(sim.p.copy())
        divs.append(sim.div.copy())
        scales.append(g.h * g.h / prm.dt_used)
        dts.append(prm.dt_used)
    label = sim.label
    rows = []
    for k in range(first, steps):
        def cycles(p0):
            return O.mg_fcg_solve(n, n, p0, divs[k], label, scales[k], 100, 1e-6)[1]
        lin = ps[k - 1] + (dts[k] / dts[k - 1]) * (ps[k - 1] - ps[k - 2])
        exact = lin + coarse_part(ps[k] - lin, label, n, n, 8)
        rows.append((cycles(ps[k - 1]), cycles(lin), cycles(exact)))
    prev, lin, exact = np.array(rows, float).mean(axis=0)
    print(f"V-cycles per solve: previous {prev:.2f}, extrapolated {lin:.2f}, "
          f"extrapolated + exact coarse correction {exact:.2f}")
    assert lin < prev
    assert lin - exact < 1.0
