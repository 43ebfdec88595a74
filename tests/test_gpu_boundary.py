"""LENNARD_JONES_INFORMATION::LJ_Force_With_PME_Direct_Force, both overloads
(Lennard_Jones_force.cuh:121-124), through the C ABI: the LJ + PME direct
force alone, from uint coordinates and charges or from the caller's
UINT_VECTOR_LJ_TYPE records with an explicit scaler and cutoff, against the
FP64 oracle's direct force."""
import sys

import numpy as np
import pytest

import oracle
from sponge_amd import system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.mark.parametrize("make", [
    lambda: system.config(0),
    lambda: system.small_random(900, (21.0, 24.5, 27.3), seed=3, cutoff=7.0, skin=1.5),
])
def test_lj_force_with_pme_direct_force_overloads(make):
    import torch
    from sponge_amd import nb
    sys_ = make()
    say(f"[case] LJ_Force_With_PME_Direct_Force {sys_.name} n={sys_.n}")
    eng = nb.NonbondedEngine(sys_, refresh_interval=0)
    eng.update_neighbor_list(forced=1)
    eng.crd_to_uint()
    ref = oracle.full_reference(sys_, beta=eng.pme.beta, K=(eng.pme.fftx, eng.pme.ffty, eng.pme.fftz))
    # overload of :124 -- adds to frc
    frc = torch.zeros_like(eng.frc)
    eng.lj.LJ_Force_With_PME_Direct_Force(eng.uint_crd, eng.charge, frc, eng.nl, eng.pme.beta)
    torch.cuda.synchronize()
    f1 = frc.cpu().numpy()
    e1 = rel_rms(f1, ref["frc_direct"])
    # it adds: a second call doubles the force
    eng.lj.LJ_Force_With_PME_Direct_Force(eng.uint_crd, eng.charge, frc, eng.nl, eng.pme.beta)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(frc.cpu().numpy(), 2.0 * f1)
    # overload of :121 -- UINT_VECTOR_LJ_TYPE records, scaler and cutoff given
    rec = torch.zeros((sys_.n, 5), dtype=torch.int32, device="cuda")
    rec[:, :3] = eng.uint_crd
    rec[:, 3] = torch.as_tensor(sys_.lj_type, dtype=torch.int32, device="cuda")
    rec[:, 4] = eng.charge.view(torch.int32)
    frc2 = torch.zeros_like(eng.frc)
    eng.lj.LJ_Force_With_PME_Direct_Force(sys_.n, rec, eng.scaler, frc2, eng.nl, sys_.cutoff, eng.pme.beta)
    torch.cuda.synchronize()
    f2 = frc2.cpu().numpy()
    say(f"    direct force rel rms: uint/charge {e1:.2e}, records {rel_rms(f2, ref['frc_direct']):.2e}")
    assert e1 < 1e-5
    # the same kernel on the same records: identical forces
    np.testing.assert_array_equal(f2, f1)
    with pytest.raises(nb.SpongeError):
        eng.lj.LJ_Force_With_PME_Direct_Force(sys_.n - 1, rec, eng.scaler, frc2, eng.nl, sys_.cutoff, eng.pme.beta)
