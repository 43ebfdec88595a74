"""sponge_lj_pme_direct_force_noalloc, the allocation-free variant of
sponge_lj_pme_direct_force (its scratch is a static device array): called
through the C ABI, it must give exactly what the allocating entry point gives
(same kernels: forces and per-atom arrays bit for bit) on two calls in a row,
the second on a smaller system, and refuse lists beyond its capacity."""
import ctypes
import sys

import numpy as np
import pytest

from sponge_amd import system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _call(entry, eng, energy):
    """One direct-force call of `entry` on fresh zeroed outputs."""
    import torch
    from sponge_amd import nb
    frc = torch.zeros_like(eng.frc)
    ae = torch.zeros_like(eng.atom_energy)
    av = torch.zeros_like(eng.atom_virial)
    ad = torch.full_like(eng.atom_direct, 7.0)
    rc = getattr(nb.load_library(), entry)(
        eng.lj.h, eng.nl.h, nb._ptr(eng.uint_crd), nb._ptr(eng.charge), nb._ptr(frc),
        ctypes.c_float(eng.pme.beta), energy, nb._ptr(ae), energy, nb._ptr(av), nb._ptr(ad), None)
    assert rc == 0, nb.load_library().sponge_nb_last_error()
    torch.cuda.synchronize()
    return [t.cpu().numpy().copy() for t in (frc, ae, av, ad)]


@pytest.mark.parametrize("energy", [0, 1])
def test_noalloc_matches_allocating_entry_point(energy):
    from sponge_amd import nb
    systems = [system.config(1), system.small_random(900, (21.0, 24.5, 27.3), seed=3, cutoff=7.0, skin=1.5)]
    for sys_ in systems:   # the second call is on a smaller input
        say(f"[case] noalloc energy={energy} {sys_.name} n={sys_.n}")
        eng = nb.NonbondedEngine(sys_, refresh_interval=0)
        eng.update_neighbor_list(forced=1)
        eng.crd_to_uint()
        got = _call("sponge_lj_pme_direct_force_noalloc", eng, energy)
        want = _call("sponge_lj_pme_direct_force", eng, energy)
        assert np.isfinite(got[0]).all()
        assert np.abs(got[0]).max() > 0
        for k, (a, b) in enumerate(zip(got, want)):
            assert np.array_equal(a, b), (sys_.name, k)


def test_noalloc_refuses_lists_beyond_its_capacity():
    """A list of more than 131,072 slots is refused before anything runs."""
    import torch
    from sponge_amd import nb
    n = 140000
    box = [112.0, 112.0, 112.0]
    say(f"[case] noalloc capacity n={n}")
    nl = nb.NEIGHBOR_LIST(n, box, 10.0, 2.0)
    lj = nb.LENNARD_JONES_INFORMATION(n, 1, np.zeros(1, np.float32), np.zeros(1, np.float32),
                                      np.zeros(n, np.int32), 10.0, box)
    dummy = torch.zeros(3 * n, dtype=torch.float32, device="cuda")
    lib = nb.load_library()
    rc = lib.sponge_lj_pme_direct_force_noalloc(lj.h, nl.h, nb._ptr(dummy), nb._ptr(dummy), nb._ptr(dummy),
                                                ctypes.c_float(0.3), 0, None, 0, None, None, None)
    assert rc == -3, rc
    assert b"static scratch" in lib.sponge_nb_last_error()
    lj.Clear()
    nl.Clear()
