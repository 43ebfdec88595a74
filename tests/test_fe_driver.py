"""The FEP / TI frame driver (sponge_amd/sponge_fe_driver): SPONGE_FEP's and
SPONGE_TI's frame loops (main_fep.cu, main_ti.cu) over SPONGE-format inputs.

CPU tests check the argument and input errors, which stop the driver before
it touches a device. The GPU test runs both programs on a three-frame
trajectory (one frame with a rescaled box) and checks the float32ene.dat
partition records, the TI average and the lambda-window file against the
FP64 oracle evaluated on the same frames.
"""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

import oracle
from sponge_amd import system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sponge_amd", "sponge_fe_driver")

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")

# partition_energy_data (FEP_core.cuh:23-40)
PARTITION = ["bond", "angle", "dihedral", "nb14_LJ", "nb14_EE", "bond_soft", "vdw_intersys", "vdw_intrasys",
             "coul_direct_intersys", "coul_direct_intrasys", "vdw_long_range_correction", "coul_long_range", "pV"]


def run(args, timeout=300):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)


@needs_driver
def test_usage_and_missing_inputs(tmp_path):
    r = run([])
    assert r.returncode == 1 and "usage" in r.stderr
    mdin = tmp_path / "mdin.txt"
    mdin.write_text("t\nframe_numbers = 2\n")
    r = run(["fep", "-mdin", str(mdin)])
    assert r.returncode == 1 and "atom numbers" in r.stderr
    mdin.write_text("t\natom_numbers = 3\n")
    r = run(["ti", "-mdin", str(mdin)])
    assert r.returncode == 1 and "trajectory" in r.stderr
    crd = tmp_path / "c.dat"
    crd.write_bytes(np.zeros(9, np.float32).tobytes())
    box = tmp_path / "b.txt"
    box.write_text("10 10 10 90 90 90\n")
    mdin.write_text(f"t\natom_numbers = 3\ncrd = {crd}\nbox = {box}\ncharge_pertubated = 1\n")
    r = run(["ti", "-mdin", str(mdin)])
    assert r.returncode == 1 and "charge A and charge B" in r.stderr


def as_read(fs):
    """The tables as the driver reads them: A/12 and B/6 parsed as FP32, then
    scaled by 12 / 6 in FP32 (LJ_soft_core.cu:1489-1513)."""
    f = lambda a, s: (np.float32(np.asarray(a, np.float64) / s) * np.float32(s)).astype(np.float32)
    return dataclasses.replace(fs, lj_A=f(fs.lj_A, 12.0), lj_B=f(fs.lj_B, 6.0), lj_A_B=f(fs.lj_A_B, 12.0),
                               lj_B_B=f(fs.lj_B_B, 6.0))


def fvol(b):
    b = np.asarray(b, np.float32)
    return float(np.float32(np.float32(b[0] * b[1]) * b[2]))


@pytest.mark.gpu
@needs_driver
def test_fep_and_ti_frames_match_oracle(tmp_path):
    from sponge_amd.system import fft_size
    fs = system.fep_mixed(system.fep_system(500, 0.35, seed=9))
    rng = np.random.default_rng(4)
    frames, boxes = [], []
    for k, s in enumerate((1.0, 1.0, 1.002)):
        b = (fs.box.astype(np.float64) * s).astype(np.float32)
        c = (fs.crd.astype(np.float64) + (rng.normal(0.0, 0.05, fs.crd.shape) if k else 0.0)) * s
        frames.append(oracle.periodic_map(c.astype(np.float32), b))
        boxes.append(b)
    nwin = 5
    mdin = system.write_fe_inputs(fs, str(tmp_path), frames, boxes, extra={"fep_lambda_windows": nwin,
                                                                          "fep_windows": tmp_path / "win.dat"})
    print(f"[case] fe driver n={fs.n} frames={len(frames)}", flush=True)
    r = run(["fep", "-mdin", mdin])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    r = run(["ti", "-mdin", mdin])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    avg = float([l for l in r.stdout.splitlines() if l.startswith("Ensemble Average")][0].split(":")[1])
    rec = np.fromfile(tmp_path / "float32ene.dat", np.float32).reshape(-1, len(PARTITION))
    assert rec.shape[0] == len(frames)
    win = np.fromfile(tmp_path / "win.dat", np.float64).reshape(len(frames), nwin, 12)

    fr = as_read(fs)
    lam0 = float(np.float32(fs.lam))
    beta = float(np.float32(oracle.pme_beta(fs.cutoff, fs.pme_tolerance)))
    K = tuple(fft_size(b) for b in fs.box)
    q = fs.charge
    dq = (fs.charge_B.astype(np.float64) - fs.charge_A.astype(np.float64)).astype(np.float32)
    p_bar = np.float32(np.float32(1.0) * np.float32(0.00001439506089041446))
    f0, f1 = oracle.lj_soft_lrc_factor(fr, 0.0), oracle.lj_soft_lrc_factor(fr, 1.0)
    dhdl = []
    for k, (c, b) in enumerate(zip(frames, boxes)):
        fk = dataclasses.replace(fr, box=b, crd=c)
        u = oracle.crd_to_uint(c, b)
        pairs = oracle.nl_brute(fk, u)
        w = oracle.fep_windows(fk, u, pairs, beta, [lam0], lambda0=lam0)[0]
        V = fvol(b)
        e_rec = oracle.pme_reciprocal(fk, u, K, beta, q)["e_rec"]
        es, en = oracle.pme_self(q, beta, b)
        e_corr = oracle.excluded(fk, u, beta, q)["e_corr"]
        ref = dict(vdw_intersys=w[0], vdw_intrasys=w[1], coul_direct_intersys=w[2], coul_direct_intrasys=w[3],
                   vdw_long_range_correction=oracle.lj_soft_lrc_factor(fr, lam0) / V,
                   coul_long_range=e_rec + es + en + e_corr,
                   pV=float(np.float32(np.float32(np.float32(p_bar * b[0]) * b[1]) * b[2])))
        scale = abs(e_rec) + abs(es + en) + abs(e_corr)
        for name, val in ref.items():
            got = float(rec[k, PARTITION.index(name)])
            tol = 2e-6 * (scale if name == "coul_long_range" else max(1.0, abs(val)))
            assert abs(got - val) <= tol, (k, name, got, val)
        for name in PARTITION[:6]:
            assert rec[k, PARTITION.index(name)] == 0.0
        # TI: LJ + direct (reference form) + long-range + PME cross terms
        t = oracle.fep_windows(fk, u, pairs, beta, [lam0], dq=dq, lambda0=lam0)[0]
        q64, d64 = q.astype(np.float64), dq.astype(np.float64)
        nf = -np.pi / (2.0 * beta ** 2 * V)
        cross = (oracle.pme_cross_reciprocal(fk, u, K, beta, q, dq)
                 - 2.0 * beta / np.sqrt(np.pi) * (q64 * d64).sum() + 2.0 * nf * q64.sum() * d64.sum()
                 + oracle.excluded_cross(fk, u, beta, q, dq))
        dhdl.append(t[4] + t[5] + (f1 - f0) / V + cross)
        # every lambda window of the frame
        lams = np.arange(nwin) / (nwin - 1)
        qw = ((1.0 - lam0) * fs.charge_A.astype(np.float64) + lam0 * fs.charge_B.astype(np.float64)).astype(np.float32)
        ww = oracle.fep_windows(fk, u, pairs, beta, lams, charge=qw, dq=dq, lambda0=lam0)
        for j, lam in enumerate(lams):
            for f in range(7):
                assert abs(win[k, j, f] - ww[j, f]) <= 1e-9 * max(1.0, abs(ww[j, f])), (k, lam, f)
            qj = (qw.astype(np.float64) + (lam - lam0) * dq.astype(np.float64)).astype(np.float32)
            pr = oracle.pme_reciprocal(fk, u, K, beta, qj)["e_rec"]
            ps = sum(oracle.pme_self(qj, beta, b))
            pc = oracle.excluded(fk, u, beta, qj)["e_corr"]
            sc = abs(pr) + abs(ps) + abs(pc)
            for f, val in enumerate((pr, ps, pc)):
                assert abs(win[k, j, 9 + f] - val) <= 1e-6 * sc, (k, lam, f, win[k, j, 9 + f], val)
    ref_avg = float(np.mean(dhdl))
    print("TI average", avg, ref_avg, flush=True)
    assert abs(avg - ref_avg) <= 1e-5 * max(1.0, abs(ref_avg)) + 1e-3
    assert open(tmp_path / "TI.txt").read().strip() == f"{avg:.6f}"
