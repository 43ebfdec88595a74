"""Pins of the oracle's pressure (oracle.virial, oracle.pressure,
oracle.barostat_run; MD_core.cu:183-186, 241-250, 1765-1795,
Lennard_Jones_force.cu:1178-1191, PME_force.cu:753-771, main.cu:334-372).

The atomic virial W of P = (2 Ek + W) / (3V) is -dU/ds for a uniform scaling
s of coordinates and box, which is how each part is pinned:
* LJ pairs: -dU_LJ/ds by central differences on a fixed set of pairs inside
  the cutoff equals sum(-f_LJ r^2) to 1e-7;
* Coulomb (Ewald with a fixed beta and grid): the total is homogeneous of
  degree -1 in the coordinates, so W = E_direct + E_correction + E_self +
  E_reciprocal; central differences of the PME energy agree to its own
  discretisation error;
* the long-range dispersion term: 6 E_lrc is the tail virial
  (numerical integration of r f(r) over r > rc against E_lrc's integrand).
The barostat loop is checked for its fixed point (mu = 1 when p = p0) and
for the Andersen piston's volume bookkeeping.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from sponge_amd import system


def scaled(sys_, s):
    return dataclasses.replace(sys_, box=(sys_.box.astype(np.float64) * s).astype(np.float32))


def test_lj_virial_is_minus_dU_ds():
    sys_ = system.config(0)
    crd = oracle.periodic_map(sys_.crd, sys_.box)
    u = oracle.crd_to_uint(crd, sys_.box)
    pairs = oracle.nl_brute(sys_, u)
    beta = oracle.pme_beta(sys_.cutoff)
    q0 = np.zeros(sys_.n, np.float32)
    h = 2.0 ** -22
    # the same uint coordinates in a scaled box are the scaled configuration;
    # the step is small enough that no pair crosses the cutoff
    e = []
    for s in (1 - h, 1 + h):
        d = oracle.direct(scaled(sys_, s), u, pairs, beta, q0)
        e.append(d)
    d1 = oracle.direct(sys_, u, pairs, beta, q0)
    assert e[0]["n_in"] == e[1]["n_in"] == d1["n_in"]
    s0 = float(np.float32(sys_.box[0] * (1 - h))) / float(sys_.box[0])
    s1 = float(np.float32(sys_.box[0] * (1 + h))) / float(sys_.box[0])
    fd = -(e[1]["e_lj"] - e[0]["e_lj"]) / (s1 - s0)
    assert abs(fd - d1["w_lj"]) < 1e-6 * abs(d1["w_lj"]), (fd, d1["w_lj"])


def test_coulomb_virial_is_the_coulomb_energy():
    """-dU/ds of the PME energy tends to E_direct + E_correction + E_self +
    E_reciprocal as the splitting and the grid are refined (the identity is
    exact for the Ewald sum; PME's discretisation error has its own scaling
    derivative, 1e-3 of the parts on the default 1 A grid)."""
    sys_ = system.config(0)
    crd = oracle.periodic_map(sys_.crd, sys_.box)
    u = oracle.crd_to_uint(crd, sys_.box)
    pairs = oracle.nl_brute(sys_, u)
    h = 2.0 ** -22
    scale = None
    gaps = []
    for tol, kf in ((1e-5, 1), (1e-9, 3)):
        beta = oracle.pme_beta(sys_.cutoff, tol)
        K = tuple(kf * system.fft_size(b) for b in sys_.box)

        def coul(s):
            ss = scaled(sys_, s)
            d = oracle.direct(ss, u, pairs, beta)
            x = oracle.excluded(ss, u, beta)
            r = oracle.pme_reciprocal(ss, u, K, beta)
            es, en = oracle.pme_self(sys_.charge, beta, ss.box)
            return (d["e_direct"] + x["e_corr"] + r["e_rec"] + es + en, d["n_in"],
                    float(ss.box[0]) / float(sys_.box[0]), abs(d["e_direct"]) + abs(x["e_corr"]))

        (um, nm, sm, _), (u0, n0, _, sc), (up, np_, sp, _) = coul(1 - h), coul(1.0), coul(1 + h)
        assert nm == n0 == np_
        scale = sc if scale is None else scale
        gaps.append(abs(-(up - um) / (sp - sm) - u0))
    assert gaps[0] < 2e-3 * scale
    assert gaps[1] < 1e-4 * scale and gaps[1] < 0.05 * gaps[0], gaps


def test_lrc_virial_is_six_times_its_energy():
    """The tail of the dispersion energy beyond rc, per unit C6 and pair
    density, is I_E = int_rc^inf -r^-6 4 pi r^2 dr; its virial integrand is
    r f(r) = -r dV/dr = 6 V(r), so the tail virial is 6 E_lrc exactly; and
    E_lrc's factor is (1/2) sum over ordered atom pairs of C6_ij I_E
    (C6 = B / 6 of the stored table)."""
    from scipy.integrate import quad
    sys_ = system.config(0)
    rc = float(sys_.cutoff)
    e_int = quad(lambda r: -1.0 / r ** 6 * 4 * np.pi * r * r, rc, np.inf)[0]
    w_int = quad(lambda r: -6.0 / r ** 6 * 4 * np.pi * r * r, rc, np.inf)[0]
    assert abs(w_int / e_int - 6.0) < 1e-10
    T = sys_.ntypes
    cnt = np.bincount(sys_.lj_type, minlength=T).astype(np.float64)
    s = 0.0
    for a in range(T):
        for b in range(T):
            hi, lo = max(a, b), min(a, b)
            s += cnt[a] * cnt[b] * float(sys_.lj_B[hi * (hi + 1) // 2 + lo]) / 6.0
    f = oracle.lj_lrc_factor(sys_)
    assert abs(f - 0.5 * s * e_int) < 1e-9 * abs(f)


def test_barostat_fixed_point_and_piston():
    sys_ = system.config(0)
    vel = system.rigid_velocities(sys_, 300.0, 5)
    pi, pj, pr = system.water_constraints(sys_)
    g = oracle.settle_groups(sys_.n, pi, pj, pr, sys_.mass)
    # one step: the Berendsen factor at the step's own pressure is 1
    _, _, _, log = oracle.barostat_run(sys_, vel, 1, 0.002, g, update_interval=1)
    p = log[0]["pressure"]
    _, _, box, log2 = oracle.barostat_run(sys_, vel, 1, 0.002, g, update_interval=1,
                                          target_pressure_bar=p / 0.00001439506089041446)
    assert abs(log2[0]["mu"] - 1.0) < 1e-12
    np.testing.assert_allclose(box, sys_.box.astype(np.float64), rtol=1e-12)
    # Andersen: V after two steps = V0 + dt (dV1 + dV2) with dV_k the piston velocity
    _, _, box, log = oracle.barostat_run(sys_, vel, 2, 0.002, g, kind="andersen", tau=0.5)
    V0 = float(np.prod(sys_.box.astype(np.float64)))
    comp = 4.5e-5 * 6.946827162543585e4
    inv_w = V0 * comp / (0.5 * 20.455) ** 2
    v1 = (log[0]["pressure"] - 0.00001439506089041446) * inv_w
    v2 = v1 + (log[1]["pressure"] - 0.00001439506089041446) * inv_w
    dt = 0.002 * 20.455
    assert abs(float(np.prod(box)) - (V0 + dt * (v1 + v2))) < 1e-9 * V0
