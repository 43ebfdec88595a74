"""The HIP leap-frog + SETTLE (include/sponge_md.h) against the FP64 oracle.

One step from the same state and the same FP32 forces: positions within
2e-5 Angstrom (a few FP32 ulps of a 20-60 Angstrom coordinate); velocities
within 1e-6 relative RMS for plain leap-frog and 5e-4 with SETTLE, whose
velocity correction is the position correction over dt, so the FP32
rounding of the unconstrained positions (up to 2e-6 Angstrom at 60
Angstrom) enters it multiplied by 1/dt = 24; the constraint pressure, whose
forces are the same corrections over dt^2, within 5e-4 relative. The
fused iteration is bit-identical to the reference's three calls (remember,
leap-frog, SETTLE). The 100-step NVE trajectory of the 648-atom rigid-water
box (the GPU's force, integrator and neighbor-list refresh against the
oracle's FP64 state) agrees within the chaos-limited bounds stated in the
test, holds every constraint to FP32 precision and conserves energy.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from sponge_amd import system

pytestmark = pytest.mark.gpu

POS_TOL = 2e-5     # Angstrom, one step
VEL_TOL = 5e-4     # relative RMS, one step with SETTLE
LF_VEL_TOL = 1e-6  # relative RMS, one plain leap-frog step
ROH, RHH = 0.9572, 2 * 0.9572 * math.sin(math.radians(104.52) / 2)
DT = 0.002 * system.TIME_UNIT_PER_PS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def min_image(d, box):
    L = np.asarray(box, np.float64)
    return d - L * np.round(d / L)


def with_pairs(sys_, n_pair_waters):
    """Constraint list where the first waters keep only their O-H1 bond (a
    rigid pair; H2 moves freely) and the rest are rigid triangles."""
    pi, pj, pr = system.water_constraints(sys_)
    keep = np.ones(len(pi), bool)
    for w in range(n_pair_waters):
        keep[3 * w + 1] = keep[3 * w + 2] = False
    return pi[keep], pj[keep], pr[keep]


def setup(sys_, constraints=None, seed=5):
    import torch
    from sponge_amd import md
    eng = md.MDEngine(sys_, refresh_interval=0, seed=seed)
    if constraints is not None:
        eng.settle = md.SETTLE(sys_.n, *constraints, sys_.mass, eng.dt)
    eng.prepare()
    eng.force()
    torch.cuda.synchronize()
    return eng


def oracle_step(sys_, eng, constraints, vmax=0.0):
    crd = eng.crd.double().cpu().numpy()
    vel = eng.vel.double().cpu().numpy()
    frc = eng.nb.frc.cpu().numpy()
    u = eng.nb.uint_crd.cpu().numpy().view(np.uint32)
    groups = oracle.settle_groups(sys_.n, *constraints, sys_.mass)
    vir = oracle.md_iteration(crd, vel, frc, sys_.mass, u, sys_.box, eng.dt, groups, vmax=vmax)
    return crd, vel, vir, groups


CASES = {
    "water648": lambda: (system.config(0), None),
    "water648_pairs": lambda: (system.config(0), 12),
    "water23570": lambda: (system.config(1), None),
}


@pytest.fixture(params=list(CASES), scope="module")
def case(request):
    sys_, npw = CASES[request.param]()
    cons = system.water_constraints(sys_) if npw is None else with_pairs(sys_, npw)
    return request.param, sys_, cons


def test_groups_match_oracle(case):
    name, sys_, cons = case
    say(f"[md] groups {name}")
    eng = setup(sys_, cons)
    g = eng.settle.groups()
    ref = oracle.settle_groups(sys_.n, *cons, sys_.mass)
    np.testing.assert_array_equal(g["tri"], ref["tri"])
    np.testing.assert_array_equal(g["pairs"], ref["pairs"])
    np.testing.assert_allclose(g["geo"], ref["geo"], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(g["plen"], ref["plen"], rtol=1e-7)
    assert g["rest"] == ref["rest"] == 0


@pytest.mark.parametrize("vmax", [0.0, 0.35])
def test_iteration_matches_oracle(case, vmax):
    import torch
    from sponge_amd import md
    name, sys_, cons = case
    say(f"[md] iteration {name} vmax {vmax}")
    eng = setup(sys_, cons)
    crd_r, vel_r, vir_r, _ = oracle_step(sys_, eng, cons, vmax)
    eng.max_velocity = vmax
    eng.pressure.zero_()
    md.md_iteration(eng.settle, eng.n, eng.vel, eng.crd, eng.nb.frc, eng.acc, eng.inverse_mass, eng.mass,
                    eng.nb.uint_crd, eng.nb.box, eng.dt, vmax, 1, eng.pressure)
    torch.cuda.synchronize()
    crd = eng.crd.double().cpu().numpy()
    vel = eng.vel.double().cpu().numpy()
    dx = np.abs(min_image(crd - crd_r, sys_.box)).max()
    dv = rel_rms(vel, vel_r)
    p = float(eng.pressure.item())
    p_ref = vir_r * 0.33333 / float(np.prod(sys_.box.astype(np.float64)))
    say(f"    max |dx| {dx:.3e} A, vel rel RMS {dv:.3e}, pressure {p:.6e} vs {p_ref:.6e}")
    assert dx < POS_TOL
    assert dv < VEL_TOL
    assert abs(p - p_ref) <= 5e-4 * abs(p_ref)
    if vmax > 0:
        speed = np.linalg.norm(vel, axis=1)
        pre = np.linalg.norm(vel_r, axis=1)
        assert speed.max() < 2 * vmax and pre.max() < 2 * vmax


def test_plain_leap_frog_matches_oracle():
    import torch
    from sponge_amd import md
    sys_ = system.config(1)
    say("[md] plain leap-frog water23570")
    eng = setup(sys_)
    crd = eng.crd.double().cpu().numpy()
    vel = eng.vel.double().cpu().numpy()
    empty = dict(tri=np.zeros((0, 3), np.int32), geo=np.zeros((0, 5)), pairs=np.zeros((0, 2), np.int32),
                 plen=np.zeros(0))
    oracle.md_iteration(crd, vel, eng.nb.frc.cpu().numpy(), sys_.mass, eng.nb.uint_crd.cpu().numpy().view(np.uint32),
                        sys_.box, eng.dt, empty)
    md.md_iteration(None, eng.n, eng.vel, eng.crd, eng.nb.frc, eng.acc, eng.inverse_mass, eng.mass, eng.nb.uint_crd,
                    eng.nb.box, eng.dt)
    torch.cuda.synchronize()
    dx = np.abs(eng.crd.double().cpu().numpy() - crd).max()
    dv = rel_rms(eng.vel.double().cpu().numpy(), vel)
    say(f"    max |dx| {dx:.3e} A, vel rel RMS {dv:.3e}")
    assert dx < POS_TOL
    assert dv < LF_VEL_TOL


def test_fused_iteration_equals_reference_calls(case):
    """sponge_md_iteration == Remember_Last_Coordinates; Leap_Frog; Do_SETTLE
    (the last two on the unconstrained atoms too), bit for bit."""
    import torch
    from sponge_amd import md
    name, sys_, cons = case
    say(f"[md] fused vs separate {name}")
    a = setup(sys_, cons)
    b = setup(sys_, cons)
    md.md_iteration(a.settle, a.n, a.vel, a.crd, a.nb.frc, a.acc, a.inverse_mass, a.mass, a.nb.uint_crd, a.nb.box,
                    a.dt)
    scaler = [x / 4294967296.0 for x in b.nb.box]
    b.settle.Remember_Last_Coordinates(b.nb.uint_crd, scaler)
    md.leap_frog(b.n, b.vel, b.crd, b.nb.frc, b.acc, b.inverse_mass, b.dt)
    b.settle.Do_SETTLE(b.mass, b.crd, b.nb.box, b.vel)
    torch.cuda.synchronize()
    assert torch.equal(a.crd, b.crd)
    assert torch.equal(a.vel, b.vel)
    assert torch.equal(a.acc, b.acc)


def test_nve_trajectory_water648():
    """100 steps of 2 fs (main.cu's loop: force, leap-frog + SETTLE, list
    refresh every 20 steps) on the GPU against the oracle's FP64 trajectory
    from the same start. The unequilibrated lattice heats from 290 K to
    ~1400 K in 0.2 ps, so the trajectories separate chaotically: the oracle
    itself, started 2e-6 Angstrom away, ends 0.003 Angstrom RMS / 0.03
    Angstrom max apart. Bounds: 1e-4 Angstrom RMS after 10 steps, 0.03
    Angstrom RMS after 100, every O-H / H-H length within 1e-4 Angstrom,
    and the total energy of both runs within 2 % of its start and of each
    other."""
    import torch
    from sponge_amd import md
    sys_ = system.config(0)
    say("[md] nve water648 100 steps")
    eng = md.MDEngine(sys_, refresh_interval=20, seed=5)
    eng.prepare()
    v0 = eng.vel.cpu().numpy().copy()
    groups = oracle.settle_groups(sys_.n, *system.water_constraints(sys_), sys_.mass)
    snaps = {}

    def rec(step, crd, vel):
        if step + 1 in (10, 100):
            snaps[step + 1] = crd.copy()

    _, _, log = oracle.md_run(sys_, v0, 100, eng.dt, groups, record=rec)
    e_gpu = []
    for k in range(100):
        ke0 = eng.kinetic_energy()
        eng.step(need_energy=1)
        pot = eng.potential_energy()
        e_gpu.append(pot + 0.5 * (ke0 + eng.kinetic_energy()))
        if k + 1 == 10:
            c10 = eng.crd.double().cpu().numpy()
    torch.cuda.synchronize()
    c100 = eng.crd.double().cpu().numpy()
    e_ref = [l["potential"] + 0.5 * (l["ke_half_before"] + l["ke_half_after"]) for l in log]
    d10 = min_image(c10 - snaps[10], sys_.box)
    d100 = min_image(c100 - snaps[100], sys_.box)
    r10, r100 = float(np.sqrt((d10 ** 2).mean())), float(np.sqrt((d100 ** 2).mean()))
    say(f"    RMS |dx| step 10 {r10:.3e}, step 100 {r100:.3e} (max {np.abs(d100).max():.3e}); "
        f"E start {e_gpu[0]:.3f} / {e_ref[0]:.3f}, end {e_gpu[-1]:.3f} / {e_ref[-1]:.3f}")
    assert r10 < 1e-4
    assert r100 < 3e-2
    x = c100.reshape(-1, 3, 3)
    oh = np.linalg.norm(min_image(x[:, 1:] - x[:, :1], sys_.box), axis=2)
    hh = np.linalg.norm(min_image(x[:, 2] - x[:, 1], sys_.box), axis=1)
    assert np.abs(oh - ROH).max() < 1e-4
    assert np.abs(hh - RHH).max() < 1e-4
    assert max(abs(e - e_gpu[0]) for e in e_gpu) < 0.02 * abs(e_gpu[0])
    assert abs(e_gpu[-1] - e_ref[-1]) < 0.02 * abs(e_ref[0])
    assert eng.nb.nl.status()["overflow"] == 0


def test_md_step_graph_replay_matches_eager():
    """Whole MD steps (force + iteration, and force + iteration + forced list
    rebuild) captured into hipGraphs and replayed for 40 steps follow the
    eager steps of a second engine from the same start (the exclusion
    correction's float atomics make the forces reproducible only to rounding,
    so the two runs agree to 1e-4 Angstrom, not bitwise)."""
    import torch
    from sponge_amd import md
    sys_ = system.config(0)
    say("[md] graph replay water648")
    a = md.MDEngine(sys_, refresh_interval=20, seed=5)
    b = md.MDEngine(sys_, refresh_interval=20, seed=5)
    for e in (a, b):
        e.prepare()
        e.step(forced_update=0)   # sizes every scratch buffer before capture
    torch.cuda.synchronize()
    g_plain, g_rebuild = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
    with torch.cuda.graph(g_plain):
        b.step(forced_update=0, stream=torch.cuda.current_stream())
    with torch.cuda.graph(g_rebuild):
        b.step(forced_update=1, stream=torch.cuda.current_stream())
    # capture records the work without running it: b is where a is
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.crd.cpu().numpy(), a.crd.cpu().numpy())
    for k in range(40):
        a.step(forced_update=1 if (k + 1) % 20 == 0 else 0)
        (g_rebuild if (k + 1) % 20 == 0 else g_plain).replay()
        if k % 10 == 9:
            torch.cuda.synchronize()
            d = min_image(a.crd.double().cpu().numpy() - b.crd.double().cpu().numpy(), sys_.box)
            assert np.abs(d).max() < 1e-4, (k, np.abs(d).max())
