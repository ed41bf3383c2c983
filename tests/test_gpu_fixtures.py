"""GPU parity on the reference's own GPU-twin fixtures, on device-side label
edits, and on the 8192^2 thermal plume (reduced for the oracle, and at full
size through size-independent properties).

* project.gputest.ts:36-75, 112-289: 40x32, Air outlet column, solid block
  i in [10,16) x j in [12,20), every field from the double-precision LCG with
  seed 98765 (u, v, p, then 3 dye planes 0.25 + 0.5|r|), DYE_KEEP
  0.9137/0.7213/1.0431, buoyancy w 0.37 dt 0.023, scale 0.02, gradScale 0.5,
  a 3-column inlet patch with zero-coverage stripes; the kernels run in the
  fused step's order (buoyancy, divergence, subtract + outflow, decay, patch)
  and every result is within 1e-6 of the FP64 CPU functions on the unrounded
  inputs.
* advect.gputest.ts:33-80, 219-273: the shear-plus-swirl fixture, both
  schemes, dye scales 1 and 2, max|diff| < 1e-5 * peak against the CPU
  advector on the unrounded inputs.
* boundaries.ts:45-64 commitLabels on the device (fh_set_labels), after a
  label edit (simulation.ts:436-439), bit-exact against the oracle.
"""
import numpy as np
import pytest

from flowlab_amd import scenes as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


# ---- project.gputest.ts -------------------------------------------------------

PNX, PNY = 40, 32
P_SCALE, P_GRAD = 0.02, 0.5
DYE_KEEP = (0.9137, 0.7213, 1.0431)
BUOY_W, BUOY_DT = 0.37, 0.023
PATCH_COLS = 3


def _project_fixture():
    """project.gputest.ts:55-87: labels and the LCG(98765) fields, drawn in
    the order u, v, p, dye (one stream)."""
    nx, ny = PNX, PNY
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[12:20, 10:16] = O.SOLID
    nu, nv, nc = (nx + 1) * ny, nx * (ny + 1), nx * ny
    r = O.lcg_fill(nu + nv + nc + 3 * nc, 98765)
    u = r[:nu].copy()
    v = r[nu:nu + nv].copy()
    p = r[nu + nv:nu + nv + nc].copy()
    dye = 0.25 + 0.5 * np.abs(r[nu + nv + nc:])
    return lab.ravel(), u, v, p, dye


def _project_patch():
    """project.gputest.ts:112-119: values vary in both directions, coverage 0
    on every 5th row and 0.4 on every 7th."""
    def fill(i, j):
        rgb = (0.1 * i + 0.01 * j, 1 - 0.02 * j, 0.5 + 0 * i)
        cov = np.where(j % 5 == 0, 0.0, np.where(j % 7 == 0, 0.4, 1.0))
        # dye.ts:46-62 leaves rgb unset (zero) where fill returns coverage 0
        z = j % 5 == 0
        rgb = tuple(np.where(z, 0.0, c) for c in rgb)
        return rgb, cov
    return S.make_dye_patch(0, 0, PATCH_COLS, PNY, fill)


def test_reference_project_fixture():
    """project.gputest.ts: buoyancy, divergence (b = -scale div), gradient
    subtraction, outflow, decay and the inlet patch, in the fused step's
    order, each within 1e-6 of the FP64 functions on the unrounded inputs."""
    from flowlab_amd.engine import HipEngine
    nx, ny = PNX, PNY
    lab, u, v, p, dye = _project_fixture()
    patch = _project_patch()
    h = 1.0 / ny
    e = HipEngine(nx, ny, 1, h)
    e.upload(u, v, p, lab, dye)
    e.set_buoyancy(BUOY_W)
    e.set_confinement(0.0)
    e.phase_forces(BUOY_DT)
    e.phase_divergence()
    b_gpu = -P_SCALE * e.read_divergence()
    e.phase_project(P_GRAD)
    gu, gv = e.read_velocity()
    # decay then the inlet: the dye pass with a zero displacement (dt = 0
    # samples every cell at its own centre) scaled by DYE_KEEP, then the patch
    e.set_dye_patch(patch)
    e.phase_dye("macCormack", 0.0, DYE_KEEP)
    gdye = e.read_dye()

    # the CPU functions, same order, on the f64 fixture
    n = nx * ny
    ov = O.apply_buoyancy(nx, ny, h, v, dye[:n], lab, BUOY_W, BUOY_DT)
    div = O.divergence(nx, ny, h, u, ov)
    b_exp = -P_SCALE * div
    ou, ov = O.subtract_gradient(nx, ny, p, u, ov, lab, P_GRAD)
    ou, ov = O.apply_outflow(nx, ny, ou, ov, lab)
    dye_exp = dye * np.repeat(np.asarray(DYE_KEEP), n)
    dye_exp = O.apply_dye_patch(nx, ny, dye_exp, 0, 0, PATCH_COLS, ny, patch.data)

    diffs = {"b": np.abs(b_gpu - b_exp).max(), "u": np.abs(gu - ou).max(),
             "v": np.abs(gv - ov).max(), "dye": np.abs(gdye - dye_exp).max()}
    print("project fixture max|diff|:", {k: f"{d:.2e}" for k, d in diffs.items()})
    for name, d in diffs.items():
        assert d < 1e-6, (name, d)
    # the fixture is not vacuous: each kernel changed its field
    assert np.abs(ov - v).max() > 1e-3 and np.abs(ou - u).max() > 1e-3
    assert np.abs(dye_exp - dye).max() > 1e-2


# ---- advect.gputest.ts --------------------------------------------------------

ANX, ANY = 40, 32


def _advect_fixture(ds):
    """advect.gputest.ts:43-80: shear plus swirl, Air outlet column, solid
    block, sharp-edged dye in world coordinates on the dye grid."""
    nx, ny = ANX, ANY
    h = 1.0 / ny
    i = np.arange(nx + 1)
    j = np.arange(ny)
    x = i[None, :] * h
    y = (j[:, None] + 0.5) * h
    u = (1.2 * np.sin(3 * y) + 0.4 * np.cos(5 * x)).ravel()
    i = np.arange(nx)
    j = np.arange(ny + 1)
    x = (i[None, :] + 0.5) * h
    y = j[:, None] * h
    v = (-0.9 * np.sin(4 * x) + 0.3 * np.cos(2 * y)).ravel()
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[12:20, 10:16] = O.SOLID
    dnx, dny, dh = nx * ds, ny * ds, h / ds
    x = (np.arange(dnx)[None, :] + 0.5) * dh
    y = (np.arange(dny)[:, None] + 0.5) * dh
    d0 = np.broadcast_to(np.where(x < 0.5 * nx * h, 1.0, 0.0), (dny, dnx))
    d1 = np.where((np.floor(x / h) + np.floor(y / h)) % 7 < 3, 0.8, 0.1)
    d2 = np.broadcast_to(np.where(y > (ny / 3) * h, 0.5, 0.0), (dny, dnx))
    dye = [np.ascontiguousarray(d, np.float64).ravel() for d in (d0, d1, d2)]
    return h, u, v, lab, dye


@pytest.mark.parametrize("ds", [1, 2])
@pytest.mark.parametrize("scheme", ["semiLagrangian", "macCormack"])
def test_reference_advect_fixture(scheme, ds):
    from flowlab_amd.engine import HipAdvector
    nx, ny = ANX, ANY
    h, u, v, lab, dye = _advect_fixture(ds)
    dt = 1.5 * h / 1.2
    adv = HipAdvector(nx, ny, ds, h)
    uo, vo = np.empty_like(u), np.empty_like(v)
    adv.velocity(None, scheme, u, v, uo, vo, lab, dt)
    qo = [np.empty_like(q) for q in dye]
    adv.dye(None, scheme, u, v, dye, qo, lab, dt)
    mac = scheme == "macCormack"
    eu, ev = O.advect_velocity(nx, ny, h, u, v, lab, dt, maccormack=mac)
    cases = [("u", eu, uo), ("v", ev, vo)]
    eq = [O.advect_scalar(nx, ny, h, u, v, dye[c], lab, dt, maccormack=mac, dscale=ds)
          for c in range(3)]
    cases.append(("dye", np.concatenate(eq), np.concatenate(qo)))
    report = []
    for name, exp, got in cases:
        peak = np.abs(exp).max()
        diff = np.abs(exp - got).max()
        assert peak > 1e-3, name
        assert diff < 1e-5 * peak, (name, diff, peak)
        report.append(f"{name} {diff / peak:.1e}")
    print(f"advect fixture {scheme} x{ds}:", " ".join(report))


# ---- commitLabels on the device ----------------------------------------------

def _label_edit(nx, ny):
    """A fluid scene with an Air outlet column, then an edit that paints a
    Solid bar and an Air patch (as a brush would) and clears part of an old
    solid block back to Fluid."""
    old = np.zeros((ny, nx), np.uint8)
    old[:, nx - 1] = O.AIR
    old[ny // 4:ny // 2, nx // 5:nx // 3] = O.SOLID
    new = old.copy()
    new[ny // 2:ny // 2 + 3, nx // 2:nx - 4] = O.SOLID
    new[2:6, 3:9] = O.AIR
    new[ny // 4:ny // 4 + 2, nx // 5:nx // 3] = O.FLUID
    new[0, 0] = O.SOLID  # a corner cell: its domain-edge faces too
    return old.ravel(), new.ravel()


@pytest.mark.parametrize("ds", [1, 2])
def test_device_commit_labels_matches_oracle(ds):
    """fh_set_labels runs boundaries.ts:45-64 on the resident fields: p zero
    off the fluid, dye zero in solids, every face touching an in-domain solid
    zero -- the same bits as the oracle's commitLabels of the same state.
    At dye scale 2 the reference's rule indexes the dye planes by the label
    index (boundaries.ts:51), which the device follows."""
    from flowlab_amd.engine import HipEngine
    nx, ny = 72, 40
    old, new = _label_edit(nx, ny)
    r = O.lcg_fill((nx + 1) * ny + nx * (ny + 1) + nx * ny, 4242)
    u = f32(r[:(nx + 1) * ny])
    v = f32(r[(nx + 1) * ny:(nx + 1) * ny + nx * (ny + 1)])
    p = f32(r[-nx * ny:])
    dnx, dny = nx * ds, ny * ds
    dye = f32(0.25 + 0.5 * np.abs(O.lcg_fill(3 * dnx * dny, 777)))
    e = HipEngine(nx, ny, ds)
    e.upload(u, v, p, old, dye)
    e.set_labels(new)
    gu, gv = e.read_velocity()
    gp = e.read_pressure()
    gd = e.read_dye()
    ep, eu, ev, ed = O.commit_labels(nx, ny, p, u, v, new, dye)
    assert np.array_equal(gu, eu) and np.array_equal(gv, ev)
    assert np.array_equal(gp, ep) and np.array_equal(gd, ed)
    # the rules themselves, on what came back
    L = new.reshape(ny, nx)
    assert not gp[new != O.FLUID].any()
    S_ = L == O.SOLID
    U = gu.reshape(ny, nx + 1)
    assert not U[:, 1:][S_].any() and not U[:, :-1][S_].any()
    V = gv.reshape(ny + 1, nx)
    assert not V[1:, :][S_].any() and not V[:-1, :][S_].any()
    if ds == 1:
        for c in range(3):
            assert not gd[c * nx * ny:(c + 1) * nx * ny][new == O.SOLID].any()
    # and the edit changed something the old labels had left alone
    assert not np.array_equal(eu, u) and not np.array_equal(ep, p)


def test_label_edit_then_steps_match_oracle():
    """Simulation.commit_label_edits (simulation.ts:436-439) mid-run: the host
    commits, the device re-labels and rebuilds its multigrid hierarchy, and
    the following steps match the oracle stepping the same edited state."""
    from flowlab_amd.simulation import Simulation
    nx, ny = 128, 64
    g, scene = S.karman_street(nx, ny)
    prm_kw = dict(scheme="macCormack", solver="mg-dc", pressure_iters=60, tol=1e-6,
                  cfl_target=0.5, residual_every=0)
    sim = Simulation(nx, ny, dict(prm_kw))
    sim.reset(scene)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    for _ in range(2):
        sim.step()
        osim.step(O.step_params(**prm_kw))
    # edit: a solid plate downstream of the cylinder and an Air hole
    L = sim.f.label.reshape(ny, nx)
    L[20:44, 80:83] = O.SOLID
    L[5:8, 100:104] = O.AIR
    # the device holds the current velocity; the host mirror is what the
    # readback returned, so the host commit acts on the same state
    sim.commit_label_edits()
    osim.label[:] = sim.f.label
    O._lib.fo_commit_labels(nx, ny, O._p(osim.p), O._p(osim.u), O._p(osim.v),
                            O._p(osim.label), O._p(osim.dye), nx * ny)
    gu, gv = sim.stepper.engine.read_velocity()
    assert rel_l2(gu, osim.u) < 1e-5 and rel_l2(gv, osim.v) < 1e-4
    for _ in range(2):
        sim.step()
        osim.step(O.step_params(**prm_kw))
        assert rel_l2(sim.f.u, osim.u) < 1e-5 and rel_l2(sim.f.v, osim.v) < 1e-4
    U = sim.f.u.reshape(ny, nx + 1)
    assert not U[20:44, 80:84].any()  # no flow through the new plate


# ---- the 8192^2 thermal plume -------------------------------------------------

def test_large_plume_reduced_matches_oracle():
    """The 8192^2 plume's generator (200 obstacles, 16 buoyant sources, the
    same SplitMix64 draws) resolved on 1024^2 with every radius scaled by 1/8:
    whole fused steps (MacCormack, buoyancy 1, MG-DC to 1e-6, dt 0.5 h)
    against the oracle's steps, every solve at 1e-6 by the oracle's FP64
    residual of the device's pressure."""
    from flowlab_amd.simulation import Simulation
    n = 1024
    g, scene = S.large_plume(n, cell_scale=n / 8192)
    kw = dict(scheme="macCormack", cfl_target=0.5, dt_max=0.5 * g.h, solver="mg-dc",
              pressure_iters=100, tol=1e-6, residual_every=0, buoyancy=1.0)
    sim = Simulation(n, n, kw)
    sim.reset(scene)
    osim = O.OracleSim(n, n)
    osim.reset(g, scene)
    pt = osim.patch
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=100, tol=1e-6,
                        cfl_target=0.5, dt_max=0.5 * g.h, buoyancy=1.0, residual_every=0,
                        patch=(pt.i0, pt.j0, pt.nx, pt.ny))
    assert (sim.f.label == O.SOLID).mean() > 0.005  # 0.8 % of the cells, as at 8192^2
    for step in range(3):
        sim.step()
        osim.step(prm, pt.data)
        e = sim.stepper.engine
        p, d = e.read_pressure(), e.read_divergence()
        rel = O.relative_residual(n, n, p, d, sim.f.label, g.h * g.h / sim.dt)
        eu, ev = rel_l2(sim.f.u, osim.u), rel_l2(sim.f.v, osim.v)
        print(f"plume 1024^2 step {step}: u {eu:.2e} v {ev:.2e} resid {rel:.2e} "
              f"cycles gpu {sim.iters} oracle {prm.iters}")
        assert rel <= 1e-6
        assert abs(sim.iters - prm.iters) <= 1
        assert eu < 1e-5 and ev < 1e-5
    dye = np.concatenate(sim.read_dye())
    assert rel_l2(dye, osim.dye) < 1e-5
    assert sim.f.v.max() > 0.0  # the sources drive an updraught


def test_large_plume_8192_properties():
    """The full 8192^2 plume (67 M cells) with the bench's solver (one MG-DC
    cycle, then flexible CG) for 3 steps: every solve meets 1e-6 by the
    oracle's FP64 residual of the device's pressure, the fields stay finite,
    no face touching a solid carries flow, the projection removes the
    divergence, and the step's max face speed is that of the read-back
    velocity."""
    from flowlab_amd.simulation import Simulation
    n = 8192
    g, scene = S.large_plume(n)
    kw = dict(scheme="macCormack", cfl_target=0.5, dt_max=0.5 * g.h, solver="mg-fcg",
              pressure_iters=100, tol=1e-6, residual_every=1, buoyancy=1.0)
    sim = Simulation(n, n, kw)
    sim.reset(scene)
    lab = sim.f.label
    fl = (lab == O.FLUID).reshape(n, n)
    Ssolid = (lab == O.SOLID).reshape(n, n)
    e = sim.stepper.engine
    for step in range(3):
        sim.step()
        p, d = e.read_pressure(), e.read_divergence()
        scale = g.h * g.h / sim.dt
        rel = O.relative_residual(n, n, p, d, lab, scale)
        assert np.isfinite(sim.f.u).all() and np.isfinite(sim.f.v).all()
        assert np.isfinite(p).all()
        U = sim.f.u.reshape(n, n + 1)
        V = sim.f.v.reshape(n + 1, n)
        assert not U[:, 1:][Ssolid].any() and not U[:, :-1][Ssolid].any()
        assert not V[1:, :][Ssolid].any() and not V[:-1, :][Ssolid].any()
        before = np.sqrt(np.mean(d.reshape(n, n)[fl] ** 2))
        after = np.sqrt(np.mean(sim.div.reshape(n, n)[fl] ** 2))
        speed = max(np.abs(sim.f.u).max(), np.abs(sim.f.v).max())
        print(f"plume 8192^2 step {step}: resid {rel:.2e} cycles {sim.iters} "
              f"div rms {before:.3e} -> {after:.3e} max speed {speed:.3e}")
        assert rel <= 1e-6
        assert after < 1e-3 * before
        assert e.max_face_speed() == speed
