"""Pins of the oracle's FEP / TI restatement (orc_fep_windows,
orc_excluded_cross; LJ_soft_core.cu:970-1418, TI_core.cu:18-125, :670-739).

Independent checks:
* at lambda = 0 and 1 the soft-core radii equal r, so the window energies are
  the plain pair energies of the A / B tables and charges (orc_direct);
* the exact dH/dlambda fields are the central finite differences in lambda of
  the window energies (LJ and Coulomb separately), with the charges moving
  along dq -- this pins every soft-core derivative term and the charge terms;
* at the module's lambda the energy totals equal orc_direct_soft's (pinned by
  its own tests) up to the soft-core-radius cutoff of the FEP kernel;
* the reference-form Coulomb derivative differs from the exact one only on
  ordinary pairs with perturbed charges (LJ_soft_core.cu:1271);
* the PME cross terms of a charge direction dq obey Euler's identity for
  quadratic forms (dq = q gives twice the energy) and polarization
  E(q + dq) - E(q) - E(dq) (charges on a dyadic grid, so sums are exact in
  FP32).
"""
import dataclasses

import numpy as np
import pytest

import oracle
from sponge_amd import system

LJ, LJIN, C, CIN, DLJ, DREF, DEX = range(7)


def setup(lam=0.35, mixed=True, n_water=300, seed=5):
    fs = system.fep_system(n_water, lam, seed=seed)
    if mixed:
        fs = system.fep_mixed(fs)
    crd = oracle.periodic_map(fs.crd.copy(), fs.box)
    u = oracle.crd_to_uint(crd, fs.box)
    pairs = oracle.nl_brute(fs, u)
    beta = oracle.pme_beta(fs.cutoff)
    dq = (fs.charge_B.astype(np.float64) - fs.charge_A.astype(np.float64)).astype(np.float32)
    return fs, u, pairs, beta, dq


def plain(fs, u, pairs, beta, state, charge):
    t, A, B = (fs.lj_type, fs.lj_A, fs.lj_B) if state == "A" else (fs.lj_type_B, fs.lj_A_B, fs.lj_B_B)
    return oracle.direct(dataclasses.replace(fs, lj_type=t, lj_A=A, lj_B=B), u, pairs, beta, charge)


def close(a, b, tol):
    return abs(a - b) <= tol * max(1.0, abs(b))


def test_end_states_are_plain():
    fs, u, pairs, beta, dq = setup()
    # windows relative to lambda0 = 0 with the A charges: lambda 1 reaches the B charges
    w = oracle.fep_windows(fs, u, pairs, beta, [0.0, 1.0], charge=fs.charge_A, dq=dq, lambda0=0.0)
    for k, (state, q) in enumerate((("A", fs.charge_A), ("B", fs.charge_B))):
        p = plain(fs, u, pairs, beta, state, q)
        assert close(w[k, LJ] + w[k, LJIN], p["e_lj"], 1e-10), (state, w[k], p["e_lj"])
        assert close(w[k, C] + w[k, CIN], p["e_direct"], 1e-10), (state, w[k], p["e_direct"])


@pytest.mark.parametrize("lam", [0.2, 0.5, 0.8])
def test_dhdl_is_the_lambda_derivative(lam):
    fs, u, pairs, beta, dq = setup()
    h = 1e-5
    # the derivative formulas are those of the energy without the FEP
    # kernel's soft-core-radius cutoff (which the dH/dlambda kernels lack)
    w = oracle.fep_windows(fs, u, pairs, beta, [lam - h, lam, lam + h], dq=dq, truncate_soft=False)
    fd_lj = (w[2, LJ] + w[2, LJIN] - w[0, LJ] - w[0, LJIN]) / (2 * h)
    fd_c = (w[2, C] + w[2, CIN] - w[0, C] - w[0, CIN]) / (2 * h)
    # FD error: O(h^2) truncation plus the rounding of the energy sums / h
    assert abs(fd_lj - w[1, DLJ]) < 1e-8 * abs(w[1, DLJ]) + 1e-7, (fd_lj, w[1, DLJ])
    assert abs(fd_c - w[1, DEX]) < 1e-8 * abs(w[1, DEX]) + 1e-7, (fd_c, w[1, DEX])
    # the derivative is not trivially small
    assert abs(w[1, DLJ]) > 1.0 and abs(w[1, DEX]) > 1.0
    # the cutoff on the soft-core radius only removes pairs within a few
    # thousandths of an Angstrom of the cutoff: a tiny change of the energy
    t = oracle.fep_windows(fs, u, pairs, beta, [lam], dq=dq)
    assert abs(t[0, C] + t[0, CIN] - w[1, C] - w[1, CIN]) < 1e-3
    assert abs(t[0, LJ] + t[0, LJIN] - w[1, LJ] - w[1, LJIN]) < 1e-3


def test_energies_match_direct_soft_at_module_lambda():
    fs, u, pairs, beta, dq = setup(mixed=False)
    # the force path's energy has no soft-core-radius cutoff
    w = oracle.fep_windows(fs, u, pairs, beta, [fs.lam], truncate_soft=False)
    d = oracle.direct_soft(fs, u, pairs, beta)
    assert close(w[0, LJ] + w[0, LJIN], d["e_lj"], 1e-10)
    assert close(w[0, C] + w[0, CIN], d["e_direct"], 1e-10)
    # solute-solvent pairs are the intersys ones, and they are all soft-core
    assert w[0, LJ] != 0.0 and w[0, C] != 0.0


def test_reference_form_differs_only_on_ordinary_pairs():
    fs, u, pairs, beta, dq = setup()
    w0 = oracle.fep_windows(fs, u, pairs, beta, [0.4])
    assert w0[0, DREF] == w0[0, DEX]
    # charges perturbed on the solute only: all its pairs are soft-core
    dq_sol = np.zeros_like(dq)
    dq_sol[fs.mask == 1] = dq[fs.mask == 1]
    w1 = oracle.fep_windows(fs, u, pairs, beta, [0.4], dq=dq_sol)
    assert close(w1[0, DREF], w1[0, DEX], 1e-12)
    w2 = oracle.fep_windows(fs, u, pairs, beta, [0.4], dq=dq)
    assert not close(w2[0, DREF], w2[0, DEX], 1e-6)


def dyadic(x, bits=10):
    return (np.round(np.asarray(x, np.float64) * 2 ** bits) / 2 ** bits).astype(np.float32)


def test_pme_cross_terms():
    fs, u, pairs, beta, dq = setup(n_water=200)
    from sponge_amd.system import fft_size
    K = tuple(fft_size(b) for b in fs.box)
    q = dyadic(fs.charge)
    d = dyadic(dq)
    rec = lambda c: oracle.pme_reciprocal(fs, u, K, beta, c)["e_rec"]
    # Euler: the derivative along q itself is twice the energy
    assert close(oracle.pme_cross_reciprocal(fs, u, K, beta, q, q), 2 * rec(q), 1e-10)
    assert close(oracle.excluded_cross(fs, u, beta, q, q), 2 * oracle.excluded(fs, u, beta, q)["e_corr"], 1e-12)
    # polarization
    s = (q.astype(np.float64) + d).astype(np.float32)
    assert close(oracle.pme_cross_reciprocal(fs, u, K, beta, q, d), rec(s) - rec(q) - rec(d), 1e-8)
    ec = lambda c: oracle.excluded(fs, u, beta, c)["e_corr"]
    assert close(oracle.excluded_cross(fs, u, beta, q, d), ec(s) - ec(q) - ec(d), 1e-10)


def test_lrc_windows_are_linear():
    fs, u, pairs, beta, dq = setup()
    f0 = oracle.lj_soft_lrc_factor(fs, 0.0)
    f1 = oracle.lj_soft_lrc_factor(fs, 1.0)
    for lam in (0.25, 0.6):
        assert close(oracle.lj_soft_lrc_factor(fs, lam), (1 - lam) * f0 + lam * f1, 1e-12)
