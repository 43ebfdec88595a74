"""Pressure and barostat coupling on the GPU against the FP64 oracle:
Calculate_Pressure_And_Potential_If_Needed's pressure (MD_core.cu:1765-1795)
from the fused step's parts and from the reference-shaped calls, and the
Berendsen and Andersen barostat loops (main.cu:334-372, 438-448) over a few
steps of the 648-atom water box.

Tolerances: the pressure within 1e-5 of (2 Ek + sum |W parts|) / (3V) (the
virial parts carry the 1e-6 energy bar, the pressure is their small net);
after 10 coupled steps the box within 1e-6 relative and the positions within
1e-4 Angstrom RMS (the NVE bound at 10 steps).
"""
import sys

import numpy as np
import pytest

import oracle
from sponge_amd import system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def min_image(d, box):
    b = np.asarray(box, np.float64)
    return d - b * np.round(d / b)


def pressure_scale(ref, ek, box):
    w = (abs(ref["w_lj"]) + 6 * abs(ref["e_lrc"]) + abs(ref["e_direct"]) + abs(ref["e_corr"]) + abs(ref["e_self"])
         + abs(ref["e_rec"]))
    return (2 * abs(ek) + w) / (3 * float(np.prod(np.asarray(box, np.float64))))


@pytest.mark.parametrize("config", [0, 1])
def test_pressure_matches_oracle(config):
    import torch
    from sponge_amd import md
    sys_ = system.config(config)
    say(f"[pressure] config {config} n={sys_.n}")
    eng = md.MDEngine(sys_, refresh_interval=0, seed=3)
    eng.prepare()
    eng.force(need_energy=0, need_virial=1)
    eng.calculate_pressure()
    torch.cuda.synchronize()
    p_fused = float(eng.pressure.cpu()[0])
    ek_gpu = float(eng.ek.cpu()[0])
    crd = eng.crd.cpu().numpy()
    vel = eng.vel.double().cpu().numpy()
    ref = oracle.full_reference(sys_, crd=crd, beta=eng.nb.pme.beta,
                                K=(eng.nb.pme.fftx, eng.nb.pme.ffty, eng.nb.pme.fftz))
    p_ref, ek = oracle.pressure(vel, sys_.mass, oracle.virial(ref), sys_.box)
    sc = pressure_scale(ref, ek, sys_.box)
    say(f"    p fused {p_fused:.6e} oracle {p_ref:.6e} scale {sc:.3e}; Ek {ek_gpu:.4f} / {ek:.4f}")
    assert abs(ek_gpu - ek) < 1e-6 * ek
    assert abs(p_fused - p_ref) < 1e-5 * sc

    # the reference-shaped calls: LJ with the atom virial, the long-range
    # correction and the PME reciprocal into d_virial, then Get_Pressure
    nb = eng.nb
    n = sys_.n
    frc = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    atom_virial = torch.zeros(n, dtype=torch.float32, device="cuda")
    atom_energy = torch.zeros(n, dtype=torch.float32, device="cuda")
    direct_atom = torch.zeros(n, dtype=torch.float32, device="cuda")
    virial = torch.zeros(1, dtype=torch.float32, device="cuda")
    pressure = torch.zeros(1, dtype=torch.float32, device="cuda")
    nb.crd_to_uint()
    nb.lj.LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(nb.nl, nb.uint_crd, nb.charge, frc, nb.pme.beta, 0,
                                                          atom_energy, 1, atom_virial, direct_atom)
    nb.lj.Long_Range_Correction(1, virial, 0, None)
    nb.pme.PME_Excluded_Force_With_Atom_Energy(nb.uint_crd, nb.scaler, nb.charge, nb.excl_start, nb.excl_list,
                                               nb.excl_numbers, frc)
    nb.pme.PME_Reciprocal_Force_With_Energy_And_Virial(nb.uint_crd, nb.charge, frc, 1, 0, virial, None,
                                                       direct_atom)
    md.get_pressure(n, eng.vel, eng.mass, atom_virial, virial, eng.volume(), pressure)
    torch.cuda.synchronize()
    p_calls = float(pressure.cpu()[0])
    say(f"    p reference-shaped calls {p_calls:.6e}")
    assert abs(p_calls - p_ref) < 1e-5 * sc


def run_barostat(kind, nsteps, **kw):
    import torch
    from sponge_amd import md
    sys_ = system.config(0)
    say(f"[barostat] {kind} water648 {nsteps} steps {kw}")
    eng = md.MDEngine(sys_, refresh_interval=0, seed=5, dt_ps=0.002)
    eng.prepare()
    v0 = eng.vel.cpu().numpy().copy()
    if kind == "andersen":
        eng.barostat = md.ANDERSEN_BAROSTAT_INFORMATION(1.0, sys_.box, tau=kw["tau"])
    else:
        eng.barostat = md.BERENDSEN_BAROSTAT_INFORMATION(1.0, sys_.box, dt_ps=0.002, tau=kw["tau"],
                                                         update_interval=kw["update_interval"])
    pressures = []
    for k in range(nsteps):
        eng.step()
        if eng.barostat.Ask_For_Calculate_Pressure(k + 1):
            pressures.append(eng.h_pressure)
    torch.cuda.synchronize()
    groups = oracle.settle_groups(sys_.n, *system.water_constraints(sys_), sys_.mass)
    crd, vel, box, log = oracle.barostat_run(sys_, v0, nsteps, 0.002, groups, kind=kind, **kw)
    return sys_, eng, pressures, crd, box, log


@pytest.mark.parametrize("kind,kw", [("berendsen", dict(tau=0.05, update_interval=5)),
                                     ("andersen", dict(tau=0.2))])
def test_barostat_matches_oracle(kind, kw):
    nsteps = 10
    sys_, eng, pressures, crd, box, log = run_barostat(kind, nsteps, **kw)
    box_gpu = np.asarray(eng.nb.box, np.float64)
    d = min_image(eng.crd.double().cpu().numpy() - crd, box)
    rms = float(np.sqrt((d ** 2).mean()))
    p_ref = [l["pressure"] for l in log]
    say(f"    box {box_gpu[0]:.7f} / {box[0]:.7f} (start {float(sys_.box[0]):.7f}); RMS |dx| {rms:.3e}; "
        f"p {pressures[-1]:.4e} / {p_ref[-1]:.4e}")
    assert len(pressures) == len(p_ref)
    assert abs(box_gpu[0] / box[0] - 1.0) < 1e-6
    assert abs(box[0] / float(sys_.box[0]) - 1.0) > 1e-5, "the barostat did not move the box"
    assert rms < 1e-4
    for a, b in zip(pressures, p_ref):
        assert abs(a - b) < 2e-3 * abs(b) + 1e-6
    status = eng.nb.nl.status()
    assert status["overflow"] == 0
