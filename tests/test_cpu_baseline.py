"""The OpenMP FP32 CPU baseline (oracle/cpu_md.cpp) computes the same forces
and trajectory as the FP64 oracle: it is timed beside the GPU as the
reference's algorithm on the host cores, so it must be that algorithm."""
import numpy as np
import pytest

import oracle
from sponge_amd import system


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.mark.parametrize("make", [
    lambda: system.config(0),
    lambda: system.small_random(900, (21.0, 24.5, 27.3), seed=3, cutoff=7.0, skin=1.5),
    lambda: system.small_random(300, (12.0, 13.0, 14.0), seed=4, cutoff=4.5, skin=1.0, fft=(12, 14, 15)),
])
def test_cpu_baseline_forces_match_oracle(make):
    s = make()
    ref = oracle.full_reference(s)
    md = oracle.CpuMd(s, np.zeros((s.n, 3), np.float32), 0.04, beta=ref["beta"], K=ref["K"])
    assert rel_rms(md.forces(), ref["frc"]) < 1e-5


def test_cpu_baseline_trajectory_matches_oracle():
    """10 NVE steps with SETTLE of the 648-atom box against the oracle's FP64
    trajectory (same forces, same leap-frog + SETTLE)."""
    s = system.config(0)
    v = system.rigid_velocities(s, 300.0, 5)
    dt = 0.002 * system.TIME_UNIT_PER_PS
    pi, pj, pr = system.water_constraints(s)
    groups = oracle.settle_groups(s.n, pi, pj, pr, s.mass)
    crd, vel, _ = oracle.md_run(s, v, 10, dt, groups)
    md = oracle.CpuMd(s, v, dt, refresh=20)
    md.run(10)
    c, vv = md.state()
    L = s.box.astype(np.float64)
    d = c - crd
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-3
    assert rel_rms(vv, vel) < 1e-3
    # rigid waters stay rigid
    o, h1 = c[0:3 * s.waters:3].astype(np.float64), c[1:3 * s.waters:3].astype(np.float64)
    r = o - h1
    r -= L * np.round(r / L)
    assert np.abs(np.linalg.norm(r, axis=1) - system.TIP3P_ROH).max() < 1e-4
