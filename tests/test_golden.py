"""Committed golden vectors of the 648-atom water box (tests/golden/).

CPU: the oracle reproduces them (a regression pin of the oracle itself; they
were made by tests/golden/make_golden.py). GPU: the HIP path through the C ABI
matches them to the parity contract (neighbor set bit-exact, forces 1e-5
relative RMS, energies 1e-6 relative).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from sponge_amd import system

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    arrays = dict(np.load(os.path.join(HERE, "water648.npz")))
    scalars = json.load(open(os.path.join(HERE, "water648.json")))
    return arrays, scalars


def test_generator_reproduces_inputs():
    arrays, s = load()
    sys_ = system.config(0)
    np.testing.assert_array_equal(sys_.crd.astype(np.float32), arrays["crd"])
    assert sys_.n == s["n"] and sys_.cutoff == s["cutoff"]


def test_oracle_reproduces_golden():
    import oracle
    arrays, s = load()
    ref = oracle.full_reference(system.config(0))
    np.testing.assert_array_equal(ref["uint"], arrays["uint"])
    pairs = np.ascontiguousarray(ref["pairs"], dtype=np.int32)
    assert len(pairs) == s["pair_count"]
    assert hashlib.sha256(pairs.tobytes()).hexdigest() == s["pair_sha256"]
    for k in ("frc", "frc_direct", "frc_excl", "frc_rec", "atom_energy", "atom_virial"):
        np.testing.assert_allclose(ref[k], arrays[k], rtol=1e-10, atol=1e-10, err_msg=k)
    for k in ("e_lj", "e_direct", "w_lj", "e_corr", "e_rec", "e_self", "e_lrc"):
        assert abs(ref[k] - s[k]) <= 1e-10 * max(1.0, abs(s[k])), k


@pytest.mark.gpu
def test_gpu_matches_golden():
    import torch
    from sponge_amd import nb
    arrays, s = load()
    sys_ = system.config(0)
    print("[golden] water648", flush=True)
    eng = nb.NonbondedEngine(sys_, refresh_interval=0)
    eng.update_neighbor_list(forced=1)
    eng.force_step(need_energy=1, need_virial=1, from_crd=True, overwrite=True)
    torch.cuda.synchronize()
    assert eng.pme.beta == np.float32(s["beta"])
    assert [eng.pme.fftx, eng.pme.ffty, eng.pme.fftz] == s["K"]
    pairs = np.ascontiguousarray(eng.nl.pairs(), dtype=np.int32)
    assert hashlib.sha256(pairs.tobytes()).hexdigest() == s["pair_sha256"]
    f = eng.frc.cpu().numpy().astype(np.float64)
    err = np.sqrt(((f - arrays["frc"]) ** 2).sum() / (arrays["frc"] ** 2).sum())
    assert err < 1e-5, err
    p = eng.energy_parts()
    for part, key in (("lj", "e_lj"), ("direct", "e_direct"), ("lj_virial", "w_lj"),
                      ("correction", "e_corr"), ("reciprocal", "e_rec"), ("self", "e_self"),
                      ("lrc", "e_lrc")):
        assert abs(p[part] - s[key]) <= 1e-6 * abs(s[key]), (part, p[part], s[key])
