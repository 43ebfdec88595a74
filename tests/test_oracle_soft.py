"""Pins of the oracle's soft-core restatement (orc_direct_soft,
orc_lj_soft_lrc_factor; LJ_soft_core.cu:159-842, 1588-1600).

Independent checks: end states reduce to the plain pair terms of the A / B
types (soft-core shifts vanish there), alpha = 0 makes the energy and force
linear in lambda between the two plain end states, finite differences of
the soft-core energy give the force, Newton's third law, and the long-range
factor reduces to the plain one at the end states.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from sponge_amd import system


def setup(lam=0.5, **kw):
    fs = system.fep_system(260, lam, seed=5)
    fs = dataclasses.replace(fs, **kw) if kw else fs
    crd = oracle.periodic_map(fs.crd.copy(), fs.box)
    u = oracle.crd_to_uint(crd, fs.box)
    pairs = oracle.nl_brute(fs, u)
    beta = oracle.pme_beta(fs.cutoff)
    return fs, u, pairs, beta


def plain(fs, u, pairs, beta, state, charge):
    t, A, B = (fs.lj_type, fs.lj_A, fs.lj_B) if state == "A" else (fs.lj_type_B, fs.lj_A_B, fs.lj_B_B)
    s = dataclasses.replace(fs, lj_type=t, lj_A=A, lj_B=B)
    return oracle.direct(s, u, pairs, beta, charge)


def close(a, b, tol=1e-10):
    return abs(a - b) <= tol * max(1.0, abs(b))


@pytest.mark.parametrize("lam,state", [(0.0, "A"), (1.0, "B")])
def test_end_states_are_plain(lam, state):
    fs, u, pairs, beta = setup(lam)
    q = fs.charge_A  # keep solute charges on at both ends to exercise the Coulomb terms
    d = oracle.direct_soft(fs, u, pairs, beta, charge=q, lam=lam)
    p = plain(fs, u, pairs, beta, state, q)
    assert d["n_in"] == p["n_in"]
    for k in ("e_lj", "e_direct", "w_lj"):
        assert close(d[k], p[k]), (k, d[k], p[k])
    np.testing.assert_allclose(d["frc"], p["frc"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(d["atom_energy"], p["atom_energy"], rtol=1e-9, atol=1e-9)


def test_alpha_zero_is_linear_in_lambda():
    lam = 0.37
    fs, u, pairs, beta = setup(lam, sc_alpha=0.0)
    q = fs.charge_A
    d = oracle.direct_soft(fs, u, pairs, beta, charge=q, lam=lam)
    a = plain(fs, u, pairs, beta, "A", q)
    b = plain(fs, u, pairs, beta, "B", q)
    assert close(d["e_lj"], (1 - lam) * a["e_lj"] + lam * b["e_lj"], 1e-9)
    assert close(d["w_lj"], (1 - lam) * a["w_lj"] + lam * b["w_lj"], 1e-9)
    assert close(d["e_direct"], a["e_direct"], 1e-9)
    np.testing.assert_allclose(d["frc"], (1 - lam) * a["frc"] + lam * b["frc"], rtol=1e-8, atol=1e-8)


def test_soft_pairs_differ_from_linear_when_alpha_on():
    """With alpha > 0 the solute-water pairs are softened: the soft-core
    energy must differ from the linear mix (guards a dropped shift)."""
    lam = 0.5
    fs, u, pairs, beta = setup(lam)
    q = fs.charge_A
    d = oracle.direct_soft(fs, u, pairs, beta, charge=q, lam=lam)
    a = plain(fs, u, pairs, beta, "A", q)
    b = plain(fs, u, pairs, beta, "B", q)
    assert abs(d["e_lj"] - ((1 - lam) * a["e_lj"] + lam * b["e_lj"])) > 1e-3


def shift_uint(u, i, d, h, box):
    step = float(box[d]) / 2 ** 32
    k = int(round(h / step))
    up, um = u.copy(), u.copy()
    up[i, d] = np.uint32((int(u[i, d]) + k) % 2 ** 32)
    um[i, d] = np.uint32((int(u[i, d]) - k) % 2 ** 32)
    return up, um, 2 * k * step


@pytest.mark.parametrize("lam", [0.25, 0.8])
def test_soft_force_is_minus_gradient(lam):
    fs, u, pairs, beta = setup(lam)
    d = oracle.direct_soft(fs, u, pairs, beta)
    nsol = int(fs.mask.sum())
    atoms = [fs.n - nsol, fs.n - 1, fs.n - nsol // 2, 0, 7]
    for i in atoms:
        for ax in range(3):
            up, um, h2 = shift_uint(u, i, ax, 1e-5, fs.box)
            ep = oracle.direct_soft(fs, up, pairs, beta)
            em = oracle.direct_soft(fs, um, pairs, beta)
            fd = -((ep["e_lj"] + ep["e_direct"]) - (em["e_lj"] + em["e_direct"])) / h2
            assert abs(fd - d["frc"][i, ax]) < 2e-4 * max(1.0, abs(d["frc"][i, ax])), (i, ax, fd, d["frc"][i, ax])


def test_newton_third_law():
    fs, u, pairs, beta = setup(0.6)
    d = oracle.direct_soft(fs, u, pairs, beta)
    assert np.abs(d["frc"].sum(0)).max() < 1e-8 * np.abs(d["frc"]).max()


def test_lrc_end_states():
    fs, _, _, _ = setup(0.0)
    a = oracle.lj_lrc_factor(fs)
    b = oracle.lj_lrc_factor(dataclasses.replace(fs, lj_type=fs.lj_type_B, lj_A=fs.lj_A_B, lj_B=fs.lj_B_B))
    assert close(oracle.lj_soft_lrc_factor(fs, 0.0), a, 1e-12)
    assert close(oracle.lj_soft_lrc_factor(fs, 1.0), b, 1e-12)
    assert close(oracle.lj_soft_lrc_factor(fs, 0.3), 0.7 * a + 0.3 * b, 1e-12)
