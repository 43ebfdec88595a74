"""Batched MD steps (sponge_batch_set_integrator / sponge_batch_md_iteration,
md.BatchedMD): every system of a batch -- soft-core FEP windows with rigid
waters and a free solute, a plain water box, and a random system without
SETTLE at its own time step and velocity cap -- integrated in one launch.

* One batched iteration gives bit for bit what sponge_md_iteration gives for
  each system from the same forces (same per-group arithmetic).
* A batched trajectory (force step + iteration + list update, rebuilds on
  the interval) follows each system's own MDEngine trajectory; rigid waters
  stay on their constraints; after the run every system's neighbor set,
  forces and energies at its current coordinates match the FP64 oracle.
* The two graphs of a batched MD step (plain, with the list rebuild)
  replay the eager steps (to the ulps of the exclusion stage's float
  atomics), with eager work between replays.
"""
import dataclasses
import math

import numpy as np
import pytest

import oracle
from sponge_amd import system

pytestmark = pytest.mark.gpu

REFRESH = 8


def say(*a):
    print(*a, flush=True)


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def make_engines():
    from sponge_amd import md
    systems = [system.fep_system(400, 0.0, seed=3), system.fep_system(400, 0.5, seed=3), system.config(0),
               system.small_random(300, (12.0, 13.0, 14.0), seed=4, cutoff=4.5, skin=1.0, fft=(12, 14, 15))]
    # the random system (unit masses, close contacts) takes a short step with
    # the velocity cap (MD_Iteration_Leap_Frog_With_Max_Velocity)
    dts = [0.002, 0.002, 0.002, 0.0002]
    vmax = [0.0, 0.0, 0.0, 2.0]
    return [md.MDEngine(s, dt_ps=dt, refresh_interval=REFRESH, seed=9 + k, max_velocity=vm)
            for k, (s, dt, vm) in enumerate(zip(systems, dts, vmax))]


def snapshot(engines):
    return [dict(crd=e.crd.clone(), vel=e.vel.clone(), frc=e.nb.frc.clone(), uint=e.nb.uint_crd.clone(),
                 acc=e.acc.clone()) for e in engines]


def restore(engines, snap):
    for e, s in zip(engines, snap):
        e.crd.copy_(s["crd"])
        e.vel.copy_(s["vel"])
        e.nb.frc.copy_(s["frc"])
        e.nb.uint_crd.copy_(s["uint"])
        e.acc.copy_(s["acc"])


def water_geometry_error(sys_, crd):
    """Largest deviation (A) of a rigid water's O-H and H-H distances from the
    constraint lengths (minimum image)."""
    nw = sys_.waters
    if not nw:
        return 0.0
    L = np.asarray(sys_.box, np.float64)
    x = np.asarray(crd, np.float64)[:3 * nw].reshape(nw, 3, 3)

    def dist(a, b):
        d = x[:, a] - x[:, b]
        d -= L * np.round(d / L)
        return np.sqrt((d * d).sum(1))
    r_hh = 2.0 * system.TIP3P_ROH * math.sin(math.radians(system.TIP3P_HOH_DEG) / 2)
    return float(max(np.abs(dist(0, 1) - system.TIP3P_ROH).max(), np.abs(dist(0, 2) - system.TIP3P_ROH).max(),
                     np.abs(dist(1, 2) - r_hh).max()))


def test_batch_iteration_matches_single_iteration():
    import torch
    from sponge_amd import md
    engines = make_engines()
    say("[case] batched iteration vs sponge_md_iteration, sizes", [e.n for e in engines])
    assert engines[3].settle is None and engines[0].settle is not None
    bmd = md.BatchedMD(engines)
    bmd.prepare()
    bmd.batch.force_step(0, 0)
    torch.cuda.synchronize()
    snap = snapshot(engines)
    bmd.iterate()
    torch.cuda.synchronize()
    got = [(e.crd.cpu().numpy().copy(), e.vel.cpu().numpy().copy(), e.acc.cpu().numpy().copy()) for e in engines]
    restore(engines, snap)
    for e in engines:
        e.iterate()
    torch.cuda.synchronize()
    for e, (x, v, a) in zip(engines, got):
        assert np.isfinite(x).all() and np.isfinite(v).all(), e.sys.name
        np.testing.assert_array_equal(x, e.crd.cpu().numpy(), err_msg=e.sys.name)
        np.testing.assert_array_equal(v, e.vel.cpu().numpy(), err_msg=e.sys.name)
        np.testing.assert_array_equal(a, e.acc.cpu().numpy(), err_msg=e.sys.name)
        assert np.abs(x - snap[engines.index(e)]["crd"].cpu().numpy()).max() > 0, e.sys.name


def test_batched_trajectory_follows_each_system_and_the_oracle():
    import torch
    from sponge_amd import md
    engines = make_engines()
    steps = 2 * REFRESH + 3
    say("[case] batched MD trajectory", steps, "steps, sizes", [e.n for e in engines])
    for e in engines:
        e.nb.nl.set_prune(0, float(e.sys.cutoff))
    start = []
    for e in engines:
        e.prepare()
        start.append((e.crd.clone(), e.vel.clone()))
    # each system on its own: MDEngine steps (list update on the interval)
    for k in range(steps):
        for e in engines:
            e.step(forced_update=1 if k % REFRESH == 0 else 0)
    torch.cuda.synchronize()
    single = [(e.crd.cpu().numpy().copy(), e.vel.cpu().numpy().copy()) for e in engines]
    for e, (x, v) in zip(engines, start):
        e.crd.copy_(x)
        e.vel.copy_(v)
    bmd = md.BatchedMD(engines)
    bmd.prepare()
    for k in range(steps):
        bmd.step(forced_update=1 if k % REFRESH == 0 else 0)
    torch.cuda.synchronize()
    for e, (x1, v1) in zip(engines, single):
        x = e.crd.cpu().numpy()
        v = e.vel.cpu().numpy()
        assert np.isfinite(x).all() and np.isfinite(v).all(), e.sys.name
        L = np.asarray(e.nb.box, np.float64)
        d = x.astype(np.float64) - x1
        d -= L * np.round(d / L)
        dx = float(np.abs(d).max())
        dv = float(np.abs(v - v1).max() / max(1e-30, np.abs(v1).max()))
        say(f"    {e.sys.name}: max |dx| {dx:.2e} A, max |dv| / max |v| {dv:.2e}")
        assert dx < 1e-3, e.sys.name
        assert dv < 1e-3, e.sys.name
        assert water_geometry_error(e.sys, x) < 1e-4, e.sys.name
    # the state reached: lists, forces and energies against the FP64 oracle
    bmd.batch.neighbor_list_update(forced=1)
    bmd.batch.force_step(1, 1)
    torch.cuda.synchronize()
    for e in engines:
        s = dataclasses.replace(e.sys, crd=e.crd.cpu().numpy().copy())
        ref = oracle.full_reference(s, beta=e.nb.pme.beta, K=(e.nb.pme.fftx, e.nb.pme.ffty, e.nb.pme.fftz))
        np.testing.assert_array_equal(e.nb.nl.pairs(), ref["pairs"], err_msg=e.sys.name)
        err = rel_rms(e.nb.frc.cpu().numpy(), ref["frc"])
        assert err < 1e-5, (e.sys.name, err)
        p = e.nb.energy_parts()
        for k, rk in (("lj", "e_lj"), ("reciprocal", "e_rec"), ("self", "e_self")):
            assert abs(p[k] - ref[rk]) <= 1e-6 * abs(ref[rk]), (e.sys.name, k, p[k], ref[rk])


def test_batched_md_graphs_replay_the_eager_steps():
    import torch
    from sponge_amd import md
    engines = make_engines()
    steps = 2 * REFRESH + 2
    say("[case] batched MD graphs,", steps, "steps")
    bmd = md.BatchedMD(engines)
    bmd.prepare()
    bmd.step(forced_update=1)                   # first use of every buffer outside a capture
    torch.cuda.synchronize()
    snap = snapshot(engines)
    for k in range(steps):
        bmd.step(forced_update=1 if k % REFRESH == 0 else 0)
    torch.cuda.synchronize()
    eager = [(e.crd.cpu().numpy().copy(), e.vel.cpu().numpy().copy()) for e in engines]
    plain, rebuild = bmd.capture()
    for rep in range(2):
        # the state of the snapshot: coordinates, velocities and the list
        # built from them (as the eager run started)
        restore(engines, snap)
        bmd.batch.neighbor_list_update(forced=1)
        torch.cuda.synchronize()
        for k in range(steps):
            (rebuild if k % REFRESH == 0 else plain).replay()
        torch.cuda.synchronize()
        # the batch's exclusion stage adds its corrections with float
        # atomics (one-sided exclusion lists, as the reference's kernel), and
        # it runs beside the reciprocal stages: the order of two corrections
        # to one atom may differ between runs by an ulp, which 18 steps carry
        # but do not amplify beyond a few ulps
        for e, (x, v) in zip(engines, eager):
            L = np.asarray(e.nb.box, np.float64)
            d = e.crd.cpu().numpy().astype(np.float64) - x
            d -= L * np.round(d / L)
            dx = float(np.abs(d).max())
            dv = rel_rms(e.vel.cpu().numpy(), v)
            say(f"    replay {rep} {e.sys.name}: max |dx| {dx:.2e} A, velocity rel rms {dv:.2e}")
            assert dx < 1e-4, (rep, e.sys.name)
            assert dv < 1e-5, (rep, e.sys.name)
        # eager work between replays
        bmd.step(forced_update=0)
        torch.cuda.synchronize()


def test_set_integrator_refuses_bad_arguments():
    from sponge_amd import md, nb
    engines = make_engines()[2:]
    bmd = md.BatchedMD(engines)
    lib = nb.load_library()
    import ctypes
    settle = (nb._P * 2)(engines[0].settle.h, engines[0].settle.h)   # 648-atom SETTLE for a 300-atom system
    io = (md._BatchMdIo * 2)()
    for i, e in enumerate(engines):
        io[i] = md._BatchMdIo(nb._ptr(e.vel), nb._ptr(e.acc), nb._ptr(e.inverse_mass), nb._ptr(e.mass),
                              float(e.dt), 0.0)
    assert lib.sponge_batch_set_integrator(bmd.batch.h, ctypes.cast(settle, nb._P), ctypes.cast(io, nb._P)) < 0
    settle[1] = None
    io[0].dt = 0.0
    assert lib.sponge_batch_set_integrator(bmd.batch.h, ctypes.cast(settle, nb._P), ctypes.cast(io, nb._P)) < 0
    # the earlier setting still works
    bmd.prepare()
    bmd.step(forced_update=1)
