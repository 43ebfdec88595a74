"""FEP / TI frame evaluators on the GPU (include/sponge_fep.h) against the
FP64 oracle: the soft-core energy split and dH/dlambda of LJ_SOFT_CORE
(LJ_soft_core.cu:970-1418, 1809-1955), every lambda window in one pass, the
charge-perturbation PME terms of TI_CORE (TI_core.cu:18-125, 670-739) and
FEP_CORE's direct-energy split (FEP_core.cu:12-66, 545-558).

The pair sums are FP64 on both sides from the same integer displacements:
1e-9 relative. PME terms built on the FP32 reciprocal grid: 1e-6 relative of
the energy scale (the force path's energy bar); the cross derivative is a
difference of such terms, 1e-5 of its own size.
"""
import sys

import numpy as np
import pytest

import oracle
from sponge_amd import system

pytestmark = pytest.mark.gpu

PAIR_TOL = 1e-9
ENERGY_TOL = 1e-6
LAMBDAS = np.linspace(0.0, 1.0, 11)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, tol, scale=None):
    s = max(1.0, abs(b)) if scale is None else scale
    return abs(a - b) <= tol * s


@pytest.fixture(scope="module")
def frame():
    import torch
    from sponge_amd import nb
    fs = system.fep_mixed(system.fep_system(900, 0.35, seed=7))
    say(f"[case] fep_mixed n={fs.n} box={fs.box.tolist()} lambda={fs.lam}")
    eng = nb.NonbondedEngine(fs, refresh_interval=0)
    eng.update_neighbor_list(forced=1)
    eng.crd_to_uint()
    torch.cuda.synchronize()
    u = eng.uint_crd.cpu().numpy().view(np.uint32)
    pairs = oracle.nl_brute(fs, u)
    dq_h = (fs.charge_B.astype(np.float64) - fs.charge_A.astype(np.float64)).astype(np.float32)
    dq = torch.as_tensor(dq_h, device="cuda")
    mask = torch.as_tensor(fs.mask, dtype=torch.int32, device="cuda")
    return fs, eng, u, pairs, dq_h, dq, mask


def lam0(fs):
    """The module's lambda_lj as it stores it (FP32, LJ_soft_core.cuh)."""
    return float(np.float32(fs.lam))


def test_energy_with_coulomb_direct(frame):
    fs, eng, u, pairs, dq_h, dq, mask = frame
    lj = eng.lj
    e_lj = lj.Get_Energy_With_Coulomb_Direct(eng.uint_crd, eng.nl, eng.charge, eng.pme.beta)
    ref = oracle.fep_windows(fs, u, pairs, eng.pme.beta, [lam0(fs)], lambda0=lam0(fs))[0]
    got = [lj.h_LJ_energy_sum_intersys, lj.h_LJ_energy_sum_intrasys, lj.h_direct_ene_sum_intersys,
           lj.h_direct_ene_sum_intrasys]
    say("energy split", got, ref[:4])
    for k in range(4):
        assert close(got[k], ref[k], PAIR_TOL), (k, got[k], ref[k])
    assert close(e_lj, ref[0] + ref[1], PAIR_TOL)


@pytest.mark.parametrize("perturbed", [0, 1])
def test_dhdl_with_coulomb_direct(frame, perturbed):
    fs, eng, u, pairs, dq_h, dq, mask = frame
    lj = eng.lj
    d_lj = lj.Get_Partial_H_Partial_Lambda_With_Columb_Direct(eng.uint_crd, eng.charge, eng.nl, dq, perturbed,
                                                              eng.pme.beta)
    ref = oracle.fep_windows(fs, u, pairs, eng.pme.beta, [lam0(fs)], dq=dq_h if perturbed else None,
                             lambda0=lam0(fs))[0]
    say("dhdl", perturbed, d_lj, lj.h_sigma_of_dH_dlambda_direct, ref[4], ref[5])
    assert close(d_lj, ref[4], PAIR_TOL)
    assert close(lj.h_sigma_of_dH_dlambda_direct, ref[5], PAIR_TOL)


def test_lambda_windows(frame):
    fs, eng, u, pairs, dq_h, dq, mask = frame
    got = eng.lj.lambda_windows(eng.nl, eng.uint_crd, eng.charge, dq, eng.pme.beta, LAMBDAS)
    ref = oracle.fep_windows(fs, u, pairs, eng.pme.beta, LAMBDAS, dq=dq_h, lambda0=lam0(fs))
    # the module's volume is the FP32 product of the box lengths (LJ_soft_core.cu:1896-1902)
    b = fs.box.astype(np.float32)
    V = float(np.float32(np.float32(b[0] * b[1]) * b[2]))
    f0, f1 = oracle.lj_soft_lrc_factor(fs, 0.0), oracle.lj_soft_lrc_factor(fs, 1.0)
    worst = 0.0
    for w, lam in enumerate(LAMBDAS):
        for k in range(7):
            err = abs(got[w, k] - ref[w, k]) / max(1.0, abs(ref[w, k]))
            worst = max(worst, err)
            assert err <= PAIR_TOL, (lam, k, got[w, k], ref[w, k])
        assert close(got[w, 7], oracle.lj_soft_lrc_factor(fs, lam) / V, 1e-12)
        assert close(got[w, 8], (f1 - f0) / V, 1e-12)
    say(f"windows: worst relative error {worst:.2e}")
    # the window at the module's lambda is the drop-in result
    one = eng.lj.lambda_windows(eng.nl, eng.uint_crd, eng.charge, dq, eng.pme.beta, [lam0(fs)])[0]
    eng.lj.Get_Energy_With_Coulomb_Direct(eng.uint_crd, eng.nl, eng.charge, eng.pme.beta)
    assert close(one[0], eng.lj.h_LJ_energy_sum_intersys, 1e-12)
    assert close(one[3], eng.lj.h_direct_ene_sum_intrasys, 1e-12)
    # long-range correction of the handle
    assert close(eng.lj.Long_Range_Correction(), oracle.lj_soft_lrc_factor(fs, lam0(fs)) / V, 1e-12)
    assert close(eng.lj.Partial_H_Partial_Lambda_Long_Range_Correction(), (f1 - f0) / V, 1e-12)


def test_pme_cross_dhdl(frame):
    fs, eng, u, pairs, dq_h, dq, mask = frame
    P = eng.pme
    K = (P.fftx, P.ffty, P.fftz)
    args = (eng.uint_crd, eng.charge, dq, eng.scaler, eng.excl_start, eng.excl_list, eng.excl_numbers)
    got = P.cross_dhdl(eng.nl, *args, lj_pertubated=0)
    q = fs.charge
    rec = oracle.pme_cross_reciprocal(fs, u, K, P.beta, q, dq_h)
    corr = oracle.excluded_cross(fs, u, P.beta, q, dq_h)
    V = float(np.prod(fs.box.astype(np.float64)))
    nf = -np.pi / (2.0 * P.beta ** 2 * V)
    q64, d64 = q.astype(np.float64), dq_h.astype(np.float64)
    slf = -2.0 * P.beta / np.sqrt(np.pi) * (q64 * d64).sum() + 2.0 * nf * q64.sum() * d64.sum()
    # the direct cross term is the exact charge derivative over all pairs with
    # the LJ unperturbed: the same walk with end state B = A and one mask
    import dataclasses
    plain = dataclasses.replace(fs, lj_type_B=fs.lj_type, mask=np.zeros_like(fs.mask))
    direct = oracle.fep_windows(plain, u, pairs, P.beta, [lam0(fs)], dq=dq_h, lambda0=lam0(fs))[0][6]
    e_rec = oracle.pme_reciprocal(fs, u, K, P.beta, q)["e_rec"]
    say("cross", got, dict(reciprocal=rec, self=slf, correction=corr, direct=direct))
    assert abs(got["reciprocal"] - rec) <= 1e-5 * abs(rec) + ENERGY_TOL * abs(e_rec)
    assert close(got["self"], slf, 1e-9)
    assert close(got["correction"], corr, PAIR_TOL)
    assert close(got["direct"], direct, PAIR_TOL)
    none = P.cross_dhdl(eng.nl, *args, lj_pertubated=1)
    assert none["direct"] == 0.0 and close(none["reciprocal"], got["reciprocal"], 1e-12)


def test_pme_lambda_windows(frame):
    fs, eng, u, pairs, dq_h, dq, mask = frame
    P = eng.pme
    K = (P.fftx, P.ffty, P.fftz)
    got = P.lambda_windows(eng.uint_crd, eng.charge, dq, eng.scaler, eng.excl_start, eng.excl_list,
                           eng.excl_numbers, fs.lam, LAMBDAS)
    for w, lam in enumerate(LAMBDAS):
        qw = (fs.charge.astype(np.float64) + (lam - lam0(fs)) * dq_h.astype(np.float64)).astype(np.float32)
        e_rec = oracle.pme_reciprocal(fs, u, K, P.beta, qw)["e_rec"]
        es, en = oracle.pme_self(qw, P.beta, fs.box)
        e_corr = oracle.excluded(fs, u, P.beta, qw)["e_corr"]
        scale = abs(e_rec) + abs(es + en) + abs(e_corr)
        for k, ref in enumerate((e_rec, es + en, e_corr)):
            assert abs(got[w, k] - ref) <= ENERGY_TOL * scale, (lam, k, got[w, k], ref)
    # lambda0 reproduces Get_Energy's parts
    g0 = P.lambda_windows(eng.uint_crd, eng.charge, None, eng.scaler, eng.excl_start, eng.excl_list,
                          eng.excl_numbers, fs.lam, [fs.lam])[0]
    a = (eng.uint_crd, eng.charge, eng.nl, eng.scaler, eng.excl_start, eng.excl_list, eng.excl_numbers)
    assert close(g0[0], P.Get_Energy(*a, which_part=P.RECIPROCAL), 1e-12)
    assert close(g0[1], P.Get_Energy(*a, which_part=P.SELF), 1e-12)
    assert close(g0[2], P.Get_Energy(*a, which_part=P.CORRECTION), 1e-12)


def test_separate_direct_energy(frame):
    fs, eng, u, pairs, dq_h, dq, mask = frame
    inter, intra = eng.pme.Seperate_Direct_Atom_Energy(eng.nl, eng.uint_crd, eng.charge, mask)
    # at lambda = 0 every soft-core radius is r: the window's Coulomb terms
    # are the plain direct energies split by mask
    ref = oracle.fep_windows(fs, u, pairs, eng.pme.beta, [0.0], lambda0=0.0)[0]
    say("separate", inter, intra, ref[2], ref[3])
    assert close(inter, ref[2], PAIR_TOL)
    assert close(intra, ref[3], PAIR_TOL)
    # no mask: everything is intrasys
    inter0, intra0 = eng.pme.Seperate_Direct_Atom_Energy(eng.nl, eng.uint_crd, eng.charge, None)
    assert inter0 == 0.0 and close(intra0, ref[2] + ref[3], PAIR_TOL)
