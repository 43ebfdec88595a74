"""Pins of the oracle's integrator and SETTLE (orc_settle_* / orc_md_iteration).

SETTLE is the closed-form solution of the SHAKE equations of a rigid
triangle: the oracle is checked against an independent Newton solve of those
equations (constraint forces along the old bond vectors, all three lengths
restored), the canonical geometry against the side lengths and centre of
mass it must reproduce, the group discovery against hand-made constraint
lists, the pair constraint against its closed form, leap-frog against the
exact constant-force recurrence, and a short NVE run against conservation of
energy and momentum.
"""
import math

import numpy as np
import pytest

import oracle
from sponge_amd import system

ROH, RHH = 0.9572, 2 * 0.9572 * math.sin(math.radians(104.52) / 2)
MO, MH = 15.999, 1.008
BIG_BOX = np.array([200.0, 200.0, 200.0], np.float32)


def _uint(x, box=BIG_BOX):
    return oracle.crd_to_uint(np.asarray(x, np.float32), box)


def _udisp(u, a, b, box=BIG_BOX):
    """Minimum-image a - b of uint coordinates (32-bit wrap-around), FP64."""
    d = (u[a].astype(np.int64) - u[b].astype(np.int64) + 2 ** 31) % 2 ** 32 - 2 ** 31
    return d * box.astype(np.float64) / 2 ** 32


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _water(rng, centre):
    th = math.radians(104.52) / 2
    body = np.array([[0, 0, 0], [ROH * math.sin(th), 0, ROH * math.cos(th)],
                     [-ROH * math.sin(th), 0, ROH * math.cos(th)]])
    return centre + body @ _rot(rng).T


@pytest.mark.parametrize("sides,masses", [
    ((ROH, ROH, RHH), (MO, MH, MH)),
    ((1.0, 1.2, 1.5), (12.0, 1.0, 14.0)),
    ((1.5, 1.0, 1.2), (1.0, 16.0, 2.0)),
])
def test_canonical_geometry(sides, masses):
    ra, rb, rc, rd, re = oracle.settle_geometry(*sides, *masses)
    A, B, C = np.array([0.0, ra]), np.array([rc, rb]), np.array([rd, re])
    mA, mB, mC = masses
    assert ra > 0
    np.testing.assert_allclose([np.linalg.norm(B - A), np.linalg.norm(C - A), np.linalg.norm(C - B)], sides,
                               rtol=1e-12)
    np.testing.assert_allclose(mA * A + mB * B + mC * C, 0.0, atol=1e-12)
    # A -> B -> C counter-clockwise
    assert (B - A)[0] * (C - A)[1] - (B - A)[1] * (C - A)[0] > 0


def test_group_discovery():
    mass = np.full(20, 1.0, np.float32)
    mass[[0, 3]] = 16.0
    # water 0-1-2 (constraints listed H-H first), water 3-4-5, a pair 6-7,
    # a chain 8-9-10-11 (degree 2 but not closed), a star 12 with 3 partners,
    # pair 14-13 (listed high index first)
    pi = [1, 0, 0, 3, 3, 4, 6, 8, 9, 10, 12, 12, 12, 14]
    pj = [2, 1, 2, 4, 5, 5, 7, 9, 10, 11, 15, 16, 17, 13]
    pr = [1.5, 1.0, 1.0, 1.0, 1.0, 1.5, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.0]
    g = oracle.settle_groups(20, pi, pj, pr, mass)
    # triangle A = lowest index, B/C = its partners in constraint order
    np.testing.assert_array_equal(g["tri"], [[0, 1, 2], [3, 4, 5]])
    np.testing.assert_array_equal(g["pairs"], [[6, 7], [13, 14]])
    np.testing.assert_allclose(g["plen"], [2.0, 3.0])
    # leftover: the chain's 3 constraints and the star's 3
    assert g["rest"] == 6
    # the first triangle's |BC| came from the 1-2 constraint
    ref = oracle.settle_geometry(1.0, 1.0, 1.5, 16.0, 1.0, 1.0)
    np.testing.assert_allclose(g["geo"][0], ref, rtol=1e-15)


def _shake(r0, r1, m, bonds, lengths):
    """Newton solve of SHAKE: r = r1 + sum_k s_k lambda_k d0_k / m with d0_k
    the old bond vectors (s = +1 on the first atom, -1 on the second), until
    every bond has its length (FP64, to 1e-14)."""
    lam = np.zeros(len(bonds))

    def positions(lam):
        r = r1.copy()
        for k, (a, b) in enumerate(bonds):
            d0 = r0[a] - r0[b]
            r[a] += lam[k] * d0 / m[a]
            r[b] -= lam[k] * d0 / m[b]
        return r

    for _ in range(50):
        r = positions(lam)
        res = np.array([np.dot(r[a] - r[b], r[a] - r[b]) - L * L for (a, b), L in zip(bonds, lengths)])
        if np.abs(res).max() < 1e-14:
            break
        J = np.zeros((len(bonds), len(bonds)))
        for i, (a, b) in enumerate(bonds):
            dr = r[a] - r[b]
            for k, (c, d) in enumerate(bonds):
                d0 = r0[c] - r0[d]
                da = (d0 / m[c] if a == c else 0) - (d0 / m[d] if a == d else 0)
                db = (d0 / m[c] if b == c else 0) - (d0 / m[d] if b == d else 0)
                J[i, k] = 2 * np.dot(dr, da - db)
        lam -= np.linalg.solve(J, res)
    return positions(lam)


@pytest.mark.parametrize("seed", range(6))
def test_settle_solves_shake(seed):
    rng = np.random.default_rng(seed)
    centre = np.array([100.0, 100.0, 100.0]) + rng.uniform(-5, 5, 3)
    r0 = _water(rng, centre)
    m = np.array([MO, MH, MH])
    # an unconstrained step: rigid motion plus 0.02-0.05 A of distortion
    r1 = r0 + rng.normal(0, 0.03, (3, 3)) + rng.normal(0, 0.05, 3)
    dt = 0.002 * system.TIME_UNIT_PER_PS
    # feed r1 through the oracle's step as positions reached with zero force
    crd = np.array(r0, np.float64)
    vel = np.array((r1 - r0) / dt, np.float64)
    u = _uint(r0)
    groups = oracle.settle_groups(3, [0, 0, 1], [1, 2, 2], [ROH, ROH, RHH], m.astype(np.float32))
    vel_before = vel.copy()
    vir = oracle.md_iteration(crd, vel, np.zeros((3, 3)), m, u, BIG_BOX, dt, groups)
    # SHAKE on the same old vectors (those of the uint coordinates)
    r0u = r0[0] + np.array([np.zeros(3), _udisp(u, 1, 0), _udisp(u, 2, 0)])
    ref = _shake(r0u, r1, m, [(0, 1), (0, 2), (1, 2)], [ROH, ROH, geo_hh(groups)])
    np.testing.assert_allclose(crd, ref, atol=1e-10)
    # lengths restored, centre of mass unchanged
    d = lambda a, b: np.linalg.norm(crd[a] - crd[b])
    np.testing.assert_allclose([d(0, 1), d(0, 2), d(1, 2)], [ROH, ROH, geo_hh(groups)], atol=1e-12)
    np.testing.assert_allclose(m @ crd / m.sum(), m @ r1 / m.sum(), atol=1e-12)
    # velocity correction = position correction / dt; virial of the
    # constraint forces F = m (r - r1) / dt^2 on the old positions
    np.testing.assert_allclose(vel - vel_before, (crd - r1) / dt, atol=1e-9)
    F = m[:, None] * (crd - r1) / dt ** 2
    np.testing.assert_allclose(vir, float((F * (r0u - r0u[0])).sum()), rtol=1e-9, atol=1e-9)


def geo_hh(groups):
    ra, rb, rc, rd, re = groups["geo"][0]
    return math.hypot(rd - rc, re - rb)


def test_settle_pair_closed_form():
    rng = np.random.default_rng(3)
    axis = np.array([1.2, 0.3, -0.4]) / np.linalg.norm([1.2, 0.3, -0.4])
    r0 = np.array([[50.0, 50.0, 50.0], [50.0, 50.0, 50.0] + 1.1 * axis])
    m = np.array([12.0, 1.5])
    dt = 0.04
    r1 = r0 + rng.normal(0, 0.04, (2, 3))
    crd, vel = r0.copy(), (r1 - r0) / dt
    u = _uint(r0)
    groups = oracle.settle_groups(2, [1], [0], [1.1], m.astype(np.float32))
    np.testing.assert_array_equal(groups["pairs"], [[0, 1]])
    vel0 = vel.copy()
    oracle.md_iteration(crd, vel, np.zeros((2, 3)), m, u, BIG_BOX, dt, groups)
    assert abs(np.linalg.norm(crd[1] - crd[0]) - np.float32(1.1)) < 1e-12
    np.testing.assert_allclose(m @ crd, m @ r1, atol=1e-12)
    # the correction of each atom is parallel to the old bond (as stored:
    # the displacement of the uint coordinates)
    old = _udisp(u, 1, 0)
    for a in range(2):
        c = np.cross(crd[a] - r1[a], old)
        assert np.linalg.norm(c) < 1e-6 * np.linalg.norm(crd[a] - r1[a]) + 1e-12
    np.testing.assert_allclose(vel - vel0, (crd - r1) / dt, atol=1e-10)


def test_leap_frog_constant_force():
    """x_n = x_0 + n dt v_0 + dt^2 a n (n + 1) / 2 for a constant force; the
    velocity cap scales v to exactly max_velocity."""
    n = 4
    m = np.array([1.0, 2.0, 16.0, 0.0])
    f = np.array([[1.0, -2.0, 0.5]] * n)
    crd = np.zeros((n, 3)) + 10.0
    vel = np.zeros((n, 3))
    vel[:, 0] = 0.1
    x0, v0 = crd.copy(), vel.copy()
    dt = 0.05
    u = _uint(crd)
    empty = dict(tri=np.zeros((0, 3), np.int32), geo=np.zeros((0, 5)), pairs=np.zeros((0, 2), np.int32),
                 plen=np.zeros(0))
    steps = 7
    for _ in range(steps):
        oracle.md_iteration(crd, vel, f, m, u, BIG_BOX, dt, empty)
    a = np.where(m > 0, 1 / np.where(m > 0, m, 1), 0)[:, None] * f
    np.testing.assert_allclose(vel, v0 + steps * dt * a, rtol=1e-13)
    np.testing.assert_allclose(crd, x0 + steps * dt * v0 + dt * dt * a * steps * (steps + 1) / 2, rtol=1e-13)
    # massless atom: zero inverse mass, pure drift
    np.testing.assert_allclose(crd[3], x0[3] + steps * dt * v0[3], rtol=1e-13)
    vel = np.array([[3.0, 4.0, 0.0]])
    crd = np.zeros((1, 3)) + 5
    oracle.md_iteration(crd, vel, np.zeros((1, 3)), np.ones(1), _uint(crd), BIG_BOX, dt, empty, vmax=2.5)
    np.testing.assert_allclose(vel, [[1.5, 2.0, 0.0]], rtol=1e-14)


def test_nve_water_conserves_energy_and_momentum():
    """The 648-atom rigid-water box, 60 steps of 2 fs from rigid-body
    velocities at 300 K: the total energy (potential + mean of the two
    half-step kinetic energies) stays within 2 % of its start while the
    unequilibrated lattice converts ~1400 kcal/mol of potential into kinetic
    energy; bond lengths stay exact and the total momentum stays at its
    (zero) start up to the PME force's small net sum."""
    s = system.config(0)
    g = oracle.settle_groups(s.n, *system.water_constraints(s), s.mass)
    assert len(g["tri"]) == 216 and len(g["pairs"]) == 0 and g["rest"] == 0
    v = system.rigid_velocities(s)
    dt = 0.002 * system.TIME_UNIT_PER_PS
    crd, vel, log = oracle.md_run(s, v, 60, dt, g)
    tot = [l["potential"] + 0.5 * (l["ke_half_before"] + l["ke_half_after"]) for l in log]
    swing = max(abs(l["potential"] - log[0]["potential"]) for l in log)
    assert swing > 1000
    assert max(abs(t - tot[0]) for t in tot) < 0.02 * abs(tot[0])
    x = crd.reshape(-1, 3, 3)
    np.testing.assert_allclose(np.linalg.norm(x[:, 1] - x[:, 0], axis=1), ROH, atol=1e-9)
    np.testing.assert_allclose(np.linalg.norm(x[:, 2] - x[:, 1], axis=1), RHH, atol=1e-6)
    m = s.mass.astype(np.float64)
    p = (m[:, None] * vel).sum(0)
    assert np.linalg.norm(p) < 1e-3 * float((m[:, None] * np.abs(vel)).sum())
