"""Batched systems (sponge_batch_*): every stage as one launch over several
independent systems of different sizes, cutoffs, PME grids and LJ modules
(soft-core FEP windows at three lambdas, a plain water box, a random box with
odd grid axes). Each system must get exactly what its own single-system calls
give (neighbor set bit-exact, forces and energy parts to FP32 atomics-order
noise) and must match the FP64 oracle to the parity contract.
"""
import numpy as np
import pytest

import oracle
from sponge_amd import system

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def make_systems():
    return [system.fep_system(400, 0.0, seed=3), system.fep_system(400, 0.5, seed=3),
            system.fep_system(400, 1.0, seed=3), system.config(0),
            system.small_random(300, (12.0, 13.0, 14.0), seed=4, cutoff=4.5, skin=1.0, fft=(12, 14, 15))]


@pytest.fixture(scope="module")
def batch_run():
    import torch
    from sponge_amd import nb
    systems = make_systems()
    print("[batch] sizes", [s.n for s in systems], flush=True)
    engines = [nb.NonbondedEngine(s, refresh_interval=0) for s in systems]
    single = []
    for e in engines:
        e.update_neighbor_list(forced=1)
        e.force_step(1, 1, from_crd=True, overwrite=True)
        torch.cuda.synchronize()
        single.append(dict(frc=e.frc.cpu().numpy().copy(), parts=e.energy_parts(), pairs=e.nl.pairs()))
    batch = nb.BatchedSystems(engines)
    for e in engines:
        e.frc.fill_(float("nan"))
        e.parts.zero_()
    batch.neighbor_list_update(forced=1)
    batch.force_step(1, 1)
    torch.cuda.synchronize()
    out = [dict(frc=e.frc.cpu().numpy().copy(), parts=e.energy_parts(), pairs=e.nl.pairs(),
                status=e.nl.status()) for e in engines]
    return systems, engines, batch, single, out


def test_batch_lists_match_single(batch_run):
    systems, engines, batch, single, out = batch_run
    for s, a, b in zip(systems, single, out):
        assert b["status"]["overflow"] == 0, s.name
        np.testing.assert_array_equal(b["pairs"], a["pairs"], err_msg=s.name)


def test_batch_forces_and_parts_match_single(batch_run):
    systems, engines, batch, single, out = batch_run
    for s, a, b in zip(systems, single, out):
        assert np.isfinite(b["frc"]).all(), s.name
        assert rel_rms(b["frc"], a["frc"]) < 1e-6, s.name
        for k in ("lj", "direct", "lj_virial", "correction", "self", "reciprocal", "lrc", "pairs_in"):
            assert abs(b["parts"][k] - a["parts"][k]) <= 1e-9 * max(1.0, abs(a["parts"][k])), (s.name, k)


def test_batch_matches_oracle(batch_run):
    systems, engines, batch, single, out = batch_run
    for s, e, b in zip(systems, engines, out):
        ref = oracle.full_reference(s, beta=e.pme.beta, K=(e.pme.fftx, e.pme.ffty, e.pme.fftz))
        np.testing.assert_array_equal(b["pairs"], ref["pairs"], err_msg=s.name)
        assert rel_rms(b["frc"], ref["frc"]) < 1e-5, s.name
        p = b["parts"]
        assert abs(p["reciprocal"] - ref["e_rec"]) <= 1e-6 * abs(ref["e_rec"]), s.name
        assert abs(p["lj"] - ref["e_lj"]) <= 1e-6 * abs(ref["e_lj"]), s.name
        assert abs(p["self"] - ref["e_self"]) <= 1e-6 * abs(ref["e_self"]), s.name


def test_batch_repeated_steps_are_stable(batch_run):  # noqa: D401
    """Skin-check updates with unmoved atoms rebuild nothing; repeated force
    steps reproduce the first."""
    import torch
    systems, engines, batch, single, out = batch_run
    before = [e.nl.status()["rebuilds"] for e in engines]
    for _ in range(3):
        batch.neighbor_list_update()
        batch.force_step(0, 0)
    torch.cuda.synchronize()
    after = [e.nl.status()["rebuilds"] for e in engines]
    assert after == before
    for s, e, b in zip(systems, engines, out):
        assert rel_rms(e.frc.cpu().numpy(), b["frc"]) < 1e-6, s.name


def test_batch_graph_replays_match_eager(batch_run):
    """One hipGraph per MD step: the captured batched step (with a forced
    list rebuild) replays to the eager result, with eager work between
    replays."""
    import torch
    systems, engines, batch, single, out = batch_run
    batch.neighbor_list_update(forced=1)
    batch.force_step(0, 0)
    torch.cuda.synchronize()
    eager = [e.frc.cpu().numpy().copy() for e in engines]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cur = torch.cuda.current_stream()
        batch.force_step(0, 0, stream=cur)
        batch.neighbor_list_update(forced=1, stream=cur)
    for r in range(3):
        for e in engines:
            e.frc.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for s, e, ref in zip(systems, engines, eager):
            assert rel_rms(e.frc.cpu().numpy(), ref) < 1e-6, (r, s.name)
        batch.force_step(0, 0)
        torch.cuda.synchronize()


def test_batch_two_stream_step_equals_serial(batch_run):
    """The force step's reciprocal stages on the batch's side stream, beside
    the pair and exclusion stages: the same forces as every stage on one
    stream, eagerly and replayed from a graph (to the ulps of the exclusion
    stage's float atomics, whose order the concurrency may change); energy
    parts to FP64 atomics-order noise."""
    import torch
    systems, engines, batch, single, out = batch_run
    batch.neighbor_list_update(forced=1)
    batch.force_step(1, 1, serial=True)
    torch.cuda.synchronize()
    ref = [(e.frc.cpu().numpy().copy(), e.energy_parts()) for e in engines]
    batch.force_step(1, 1)
    torch.cuda.synchronize()
    for s, e, (f, p) in zip(systems, engines, ref):
        assert rel_rms(e.frc.cpu().numpy(), f) < 1e-7, s.name
        for k in ("lj", "direct", "correction", "self", "reciprocal", "lrc"):
            assert abs(e.energy_parts()[k] - p[k]) <= 1e-9 * max(1.0, abs(p[k])), (s.name, k)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batch.force_step(0, 0, stream=torch.cuda.current_stream())
    for r in range(2):
        for e in engines:
            e.frc.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for s, e, (f, p) in zip(systems, engines, ref):
            assert rel_rms(e.frc.cpu().numpy(), f) < 1e-7, (r, s.name)
        batch.force_step(0, 0, serial=True)
        torch.cuda.synchronize()


def test_config4_batch_matches_oracle():
    """The batched-FEP workload the benchmark times: 16 soft-core lambda
    windows of a solvated 12-atom solute and 8 replicas of 1.4k-12k atoms
    (24 systems, 100,875 atoms, cutoff 9 A, per-system PME grids), stepped as
    one batch (one launch per stage, forced list rebuild, energy step). Every
    system's neighbor set, forces and energy parts match the FP64 oracle."""
    import torch
    from sponge_amd import nb
    systems = system.fep_batch()
    print(f"[case] config-4 batch: {len(systems)} systems, {sum(s.n for s in systems)} atoms", flush=True)
    engines = [nb.NonbondedEngine(s, refresh_interval=0) for s in systems]
    batch = nb.BatchedSystems(engines)
    batch.neighbor_list_update(forced=1)
    batch.force_step(1, 1)
    torch.cuda.synchronize()
    worst = 0.0
    for s, e in zip(systems, engines):
        assert e.nl.status()["ok"], s.name
        ref = oracle.full_reference(s, beta=e.pme.beta, K=(e.pme.fftx, e.pme.ffty, e.pme.fftz))
        np.testing.assert_array_equal(e.nl.pairs(), ref["pairs"], err_msg=s.name)
        err = rel_rms(e.frc.cpu().numpy(), ref["frc"])
        worst = max(worst, err)
        assert err < 1e-5, (s.name, err)
        p = e.energy_parts()
        for k, rk in (("lj", "e_lj"), ("reciprocal", "e_rec"), ("self", "e_self"), ("correction", "e_corr")):
            assert abs(p[k] - ref[rk]) <= 1e-6 * abs(ref[rk]), (s.name, k, p[k], ref[rk])
        assert p["pairs_in"] == ref["n_in"], s.name
    print(f"    worst force rel rms {worst:.2e}", flush=True)
