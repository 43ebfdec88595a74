"""Pins the FP64 oracle before it is trusted as the checker.

The reference ships no tests, golden outputs or expected energies, so every
check here comes from closed forms, invariants, brute force or an independent
library routine (numpy's pocketfft), each chosen so that a plausible mistake
(a dropped term, a wrong sign, a transposed index, a factor 2) fails.
"""
import math

import numpy as np
import pytest

import oracle
from sponge_amd import system

K_E = 18.2223 ** 2  # kcal A / (mol e^2) in SPONGE's charge units


def two_atom_system(r, q=(0.0, 0.0), types=(0, 0), sigma=3.0, eps=0.2, box=40.0):
    A = np.array([12 * 4 * eps * sigma ** 12], np.float32)
    B = np.array([6 * 4 * eps * sigma ** 6], np.float32)
    crd = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]], np.float32)
    return system.System(
        name="dimer", box=np.full(3, box, np.float32), crd=crd,
        charge=np.array(q, np.float32) * np.float32(system.CHARGE_SCALE),
        mass=np.ones(2, np.float32), lj_type=np.array(types, np.int32), lj_A=A, lj_B=B,
        excl_start=np.zeros(2, np.int32), excl_numbers=np.zeros(2, np.int32),
        excl_list=np.zeros(0, np.int32), cutoff=10.0, skin=2.0)


def shift_uint(u, i, d, h, box):
    """u with atom i moved by about +-h along d in whole uint steps; returns
    (u_plus, u_minus, exact 2h in Angstrom)."""
    step = float(box[d]) / 2 ** 32
    k = int(round(h / step))
    up, um = u.copy(), u.copy()
    up[i, d] = np.uint32((int(u[i, d]) + k) % 2 ** 32)
    um[i, d] = np.uint32((int(u[i, d]) - k) % 2 ** 32)
    return up, um, 2 * k * step


# ---------------------------------------------------------------- coordinates
def test_uint_closed_form_and_wrap():
    box = np.array([18.64, 18.64, 18.64], np.float32)
    x = np.array([[0.0, 9.32, 18.63]], np.float32)
    u = oracle.crd_to_uint(x, box)[0]
    assert u[0] == 0
    cof = np.float32(1073741824.0) / box[0]
    assert u[1] == (int(np.float32(9.32) * cof) << 2)
    # wrap invariance: x and x - L map to the same fixed point up to one LSB
    # (truncation toward zero). Box 16 and x in [8, 16) keep x - L and x * cof
    # exact in FP32, so only the truncation can differ.
    box16 = np.full(3, 16.0, np.float32)
    rng = np.random.default_rng(0)
    y = rng.uniform(8, 16, (500, 3)).astype(np.float32)
    a = oracle.crd_to_uint(y, box16).astype(np.int64)
    b = oracle.crd_to_uint(y - box16, box16).astype(np.int64)
    d = ((a - b + 2 ** 31) % 2 ** 32) - 2 ** 31
    assert set(np.unique(d).tolist()) <= {0, 4}


def test_periodic_map_one_box():
    box = np.array([10.0, 10.0, 10.0], np.float32)
    c = np.array([[-1.0, 10.0, 5.0], [12.0, -0.5, 9.99]], np.float32)
    m = oracle.periodic_map(c, box)
    np.testing.assert_allclose(m, [[9.0, 0.0, 5.0], [2.0, 9.5, 9.99]], rtol=0, atol=1e-6)


# ------------------------------------------------------------- neighbor list
@pytest.mark.parametrize("case", ["cfg0_3cells", "random_4cells", "random_5cells", "noncubic"])
def test_nl_reference_grid_equals_brute_force(case):
    if case == "cfg0_3cells":
        s = system.config(0)
        want_dims = (3, 3, 3)
    elif case == "random_4cells":
        s = system.small_random(500, (19.0, 19.0, 19.0), seed=1, cutoff=7.0, skin=2.0)
        want_dims = (4, 4, 4)
    elif case == "random_5cells":
        s = system.small_random(700, (26.0, 26.0, 26.0), seed=2, cutoff=8.0, skin=2.0)
        want_dims = (5, 5, 5)
    else:
        s = system.small_random(900, (21.0, 24.5, 33.3), seed=3, cutoff=7.0, skin=1.5)
        want_dims = None
    crd = oracle.periodic_map(s.crd, s.box)
    u = oracle.crd_to_uint(crd, s.box)
    brute = oracle.nl_brute(s, u)
    grid, dims = oracle.nl_grid(s, s.crd)
    if want_dims:
        assert dims == want_dims
    # the reference's stencil never visits a cell pair twice: no duplicates
    assert len(np.unique(grid, axis=0)) == len(grid)
    np.testing.assert_array_equal(grid, brute)
    assert (brute[:, 0] < brute[:, 1]).all()


def test_nl_excludes_exclusions_and_counts():
    s = system.config(0)
    crd = oracle.periodic_map(s.crd, s.box)
    u = oracle.crd_to_uint(crd, s.box)
    p = oracle.nl_brute(s, u)
    keys = set(map(tuple, p.tolist()))
    for w in range(s.waters):
        o = 3 * w
        for a, b in ((o, o + 1), (o, o + 2), (o + 1, o + 2)):
            assert (a, b) not in keys
    # brute force over every pair with the same FP32 test, without exclusions
    cof = (np.float32(1.0) / np.float32(4294967296.0)) * s.box
    d = (u[None, :, :].astype(np.int64) - u[:, None, :].astype(np.int64))
    d = (((d + 2 ** 31) % 2 ** 32) - 2 ** 31).astype(np.int32).astype(np.float32) * cof
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    rl = np.float32(s.cutoff) + np.float32(s.skin)
    within = np.triu(r2 < rl * rl, 1)
    assert within.sum() == len(p) + 3 * s.waters


# -------------------------------------------------------------- direct space
def test_lj_dimer_closed_forms():
    """E(r) = 4 eps ((s/r)^12 - (s/r)^6) at the separation the uint
    coordinates encode (FP32 crd -> uint rounds positions by ~1e-6 A):
    zero at r = sigma, -eps with zero force at r = 2^(1/6) sigma."""
    sigma, eps = 3.0, 0.2
    for r, e_want, f_zero in ((sigma, 0.0, False), (2 ** (1 / 6) * sigma, -eps, True)):
        s = two_atom_system(r, sigma=sigma, eps=eps)
        u = oracle.crd_to_uint(s.crd, s.box)
        d = oracle.direct(s, u, np.array([[0, 1]], np.int32), beta=0.3)
        ru = float(u[1, 0].astype(np.int64) - u[0, 0]) * float(s.box[0]) / 2 ** 32
        A0, B0 = float(s.lj_A[0]) / 12, float(s.lj_B[0]) / 6
        assert d["e_lj"] == pytest.approx(A0 / ru ** 12 - B0 / ru ** 6, rel=1e-12, abs=1e-12)
        assert d["e_lj"] == pytest.approx(e_want, abs=1e-5)
        if f_zero:
            assert abs(d["frc"][0, 0]) < 1e-4
    # LJ force is minus the derivative of the energy (central difference)
    s0 = two_atom_system(3.5, sigma=sigma, eps=eps)
    u = oracle.crd_to_uint(s0.crd, s0.box)
    pr = np.array([[0, 1]], np.int32)
    f = oracle.direct(s0, u, pr, beta=0.3)["frc"][1, 0]
    up, um, hh = shift_uint(u, 1, 0, 1e-4, s0.box)
    e = [oracle.direct(s0, uu, pr, 0.3)["e_lj"] for uu in (up, um)]
    assert f == pytest.approx(-(e[0] - e[1]) / hh, rel=1e-7)


def test_coulomb_pair_closed_form():
    beta, r = 0.3, 4.0
    s = two_atom_system(r, q=(0.8, -0.5), types=(0, 0), sigma=0.0, eps=0.0)
    s.lj_A[:] = 0
    s.lj_B[:] = 0
    u = oracle.crd_to_uint(s.crd, s.box)
    d = oracle.direct(s, u, np.array([[0, 1]], np.int32), beta)
    qq = float(s.charge[0]) * float(s.charge[1])
    rr = (u[1, 0].astype(np.int64) - u[0, 0]) * (s.box[0] / 2 ** 32)
    assert d["e_direct"] == pytest.approx(qq * math.erfc(beta * rr) / rr, rel=1e-12)
    # attractive pair: force on atom 1 points to atom 0 (-x)
    assert d["frc"][1, 0] < 0
    assert d["frc"][0, 0] == pytest.approx(-d["frc"][1, 0], rel=1e-14)
    dE = qq * (-(2 * beta / math.sqrt(math.pi)) * math.exp(-(beta * rr) ** 2) * rr - math.erfc(beta * rr)) / rr ** 2
    assert d["frc"][1, 0] == pytest.approx(-dE, rel=1e-10)


def test_excluded_pair_closed_form():
    """direct + correction of one excluded pair is -qq erf(b r)/r + qq erfc/r
    = qq (1 - 2 erf)/r ... i.e. correction alone is -qq erf(b r)/r and its
    force is minus the derivative."""
    beta, r = 0.35, 0.9572
    s = two_atom_system(r, q=(-0.834, 0.417))
    s.excl_numbers[:] = [1, 0]
    s.excl_start[:] = [0, 1]
    s.excl_list = np.array([1], np.int32)
    u = oracle.crd_to_uint(s.crd, s.box)
    x = oracle.excluded(s, u, beta)
    qq = float(s.charge[0]) * float(s.charge[1])
    rr = (u[1, 0].astype(np.int64) - u[0, 0]) * (s.box[0] / 2 ** 32)
    assert x["e_corr"] == pytest.approx(-qq * math.erf(beta * rr) / rr, rel=1e-12)
    g = lambda t: -qq * math.erf(beta * t) / t
    h = 1e-5
    assert x["frc"][1, 0] == pytest.approx(-(g(rr + h) - g(rr - h)) / (2 * h), rel=1e-7)


def test_direct_forces_match_finite_differences():
    s = system.small_random(60, (14.0, 15.0, 16.0), seed=11, cutoff=6.0, skin=1.0)
    beta = oracle.pme_beta(s.cutoff)
    crd = oracle.periodic_map(s.crd, s.box)
    u = oracle.crd_to_uint(crd, s.box)
    pairs = oracle.nl_brute(s, u)
    F = oracle.direct(s, u, pairs, beta)["frc"] + oracle.excluded(s, u, beta)["frc"]
    rng = np.random.default_rng(1)
    for _ in range(6):
        i, d = int(rng.integers(s.n)), int(rng.integers(3))
        up, um, hh = shift_uint(u, i, d, 1e-5, s.box)
        e = []
        for uu in (up, um):
            ee = oracle.direct(s, uu, pairs, beta)
            e.append(ee["e_lj"] + ee["e_direct"] + oracle.excluded(s, uu, beta)["e_corr"])
        fd = -(e[0] - e[1]) / hh
        roundoff = 64 * np.finfo(float).eps * max(abs(e[0]), abs(e[1])) / hh
        assert F[i, d] == pytest.approx(fd, rel=1e-6, abs=roundoff)


# ---------------------------------------------------------------- PME setup
def test_beta_and_grid_sizes_match_reference_rules():
    assert oracle.pme_beta(8.0) == pytest.approx(0.348639, abs=2e-6)
    assert oracle.pme_beta(10.0) == pytest.approx(0.275106, abs=2e-6)
    assert oracle.pme_beta(9.0) == pytest.approx(0.307676, abs=2e-6)
    b = oracle.pme_beta(10.0)
    assert math.erfc(b * 10.0) / 10.0 == pytest.approx(1e-5, rel=1e-6)
    assert [system.fft_size(x) for x in (18.64, 61.9, 98.6)] == [20, 64, 100]
    assert system.fft_size(125.0) == 128    # snapped
    assert system.fft_size(40.1) == 44 or system.check_2357(system.fft_size(40.1))
    for L in np.linspace(10, 200, 57):
        k = system.fft_size(L)
        assert system.check_2357(k) and k % 4 == 0
        assert k >= L or k in (64, 128, 256, 512, 1024)   # snapped sizes may fall below L


def test_bspline_moduli():
    assert oracle.bspline_mod2(0, 64) == pytest.approx(1.0, abs=1e-14)
    # order 4 at k = K/2: |1/6 - 4/6 + 1/6|^-2 = 9
    assert oracle.bspline_mod2(32, 64) == pytest.approx(9.0, rel=1e-12)
    for k in range(1, 20):
        assert oracle.bspline_mod2(k, 20) == pytest.approx(oracle.bspline_mod2(20 - k, 20), rel=1e-12)


def test_bc_table_zero_mode_and_symmetry():
    box = np.array([20.0, 22.0, 24.0], np.float32)
    bc = oracle.pme_bc_full((12, 14, 15), box, 0.35)
    assert bc[0, 0, 0] == 0
    assert bc[1, 2, 3] == pytest.approx(bc[11, 12, 12], rel=1e-14)
    assert (bc >= 0).all()


# --------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(20, 20, 20), (12, 14, 15), (8, 9, 10), (6, 7, 21)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(5)
    a = rng.normal(size=shape)
    c = oracle.fft3_r2c(a)
    np.testing.assert_allclose(c, np.fft.rfftn(a), rtol=0, atol=1e-9 * np.abs(a).sum())
    b = oracle.fft3_c2r(c, shape[2])
    np.testing.assert_allclose(b, a * np.prod(shape), rtol=0, atol=1e-8 * np.prod(shape))


def test_fft_delta_and_plane_wave():
    a = np.zeros((8, 8, 8))
    a[0, 0, 0] = 1
    np.testing.assert_allclose(oracle.fft3_r2c(a), np.ones((8, 8, 5)), atol=1e-14)
    x = np.arange(8)
    a = np.cos(2 * np.pi * 3 * x / 8)[None, None, :] * np.ones((8, 8, 1))
    c = oracle.fft3_r2c(a)
    nz = np.argwhere(np.abs(c) > 1e-9)
    assert nz.tolist() == [[0, 0, 3]]
    assert c[0, 0, 3].real == pytest.approx(8 * 8 * 8 / 2)


# --------------------------------------------------------------- reciprocal
def _neutral_points(n, box, seed):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, 1, (n, 3)) * box
    q = rng.normal(0, 1, n)
    q -= q.mean()
    return xyz, q


def _spme_total(xyz, q, box, beta, K, rc):
    """direct (all minimum-image pairs within rc) + SPME reciprocal + self."""
    n = len(q)
    s = system.System(name="pts", box=np.asarray(box, np.float32), crd=xyz.astype(np.float32),
                      charge=np.asarray(q, np.float32), mass=np.ones(n, np.float32),
                      lj_type=np.zeros(n, np.int32), lj_A=np.zeros(1, np.float32),
                      lj_B=np.zeros(1, np.float32), excl_start=np.zeros(n, np.int32),
                      excl_numbers=np.zeros(n, np.int32), excl_list=np.zeros(0, np.int32),
                      cutoff=rc, skin=0.0)
    u = oracle.crd_to_uint(oracle.periodic_map(s.crd, s.box), s.box)
    pairs = np.array([(i, j) for i in range(n) for j in range(i + 1, n)], np.int32)
    d = oracle.direct(s, u, pairs, beta)
    r = oracle.pme_reciprocal(s, u, K, beta)
    es, en = oracle.pme_self(s.charge, beta, s.box)
    xyz_q = u.astype(np.float64) * (np.asarray(box, np.float64) / 2 ** 32)
    return d["e_direct"] + r["e_rec"] + es + en, d["frc"] + r["frc"], xyz_q, s.charge.astype(np.float64), r


def test_spme_converges_to_exact_ewald():
    box = np.array([12.0, 12.0, 12.0])
    xyz, q = _neutral_points(10, box, seed=3)
    q = q * 5.0
    beta, rc = 0.45, 5.9
    errs = []
    for K in (24, 48, 96):
        e_spme, f_spme, xyz_q, qq, _ = _spme_total(xyz, q, box, beta, (K, K, K), rc)
        e_ex, f_ex = oracle.ewald_exact(xyz_q, qq, box.astype(np.float32), beta, rc, 12)
        errs.append((abs(e_spme - e_ex) / abs(e_ex),
                     np.sqrt(((f_spme - f_ex) ** 2).sum() / (f_ex ** 2).sum())))
    # order-4 SPME converges to the exact Ewald sum as the grid refines
    assert errs[1][0] < errs[0][0] / 8 and errs[2][0] < errs[1][0] / 8
    assert errs[2][0] < 1e-6 and errs[2][1] < 3e-5


def test_ewald_total_independent_of_beta():
    box = np.array([10.0, 11.0, 12.0])
    xyz, q = _neutral_points(6, box, seed=4)
    e1, _ = oracle.ewald_exact(xyz, q, box.astype(np.float32), 0.35, 30.0, 14)
    e2, _ = oracle.ewald_exact(xyz, q, box.astype(np.float32), 0.5, 30.0, 18)
    assert e1 == pytest.approx(e2, rel=1e-9)


def test_nacl_madelung_constant():
    """Rock salt, 3x3x3 conventional cells of edge a: E / N_pairs = -M q^2/(a/2)."""
    a, ncell = 5.64, 3
    pts, qs = [], []
    for i in range(2 * ncell):
        for j in range(2 * ncell):
            for k in range(2 * ncell):
                pts.append([i * a / 2, j * a / 2, k * a / 2])
                qs.append(1.0 if (i + j + k) % 2 == 0 else -1.0)
    xyz = np.array(pts) + 0.1
    q = np.array(qs) * system.CHARGE_SCALE
    box = np.full(3, a * ncell)
    beta, rc = 0.45, 8.4
    e_spme, f, xyz_q, qq, _ = _spme_total(xyz, q, box, beta, (96, 96, 96), rc)
    M = 1.747564594633
    npairs = len(q) / 2
    want = -npairs * M * K_E / (a / 2)
    assert e_spme == pytest.approx(want, rel=1e-7)
    assert np.abs(f).max() < 1e-3   # every ion sits at a centre of symmetry


def test_reciprocal_force_is_gradient():
    box = np.array([13.0, 14.0, 15.0])
    xyz, q = _neutral_points(8, box, seed=9)
    q = q * 10
    beta, K = 0.4, (16, 18, 20)
    _, _, xyz_q, qq, r = _spme_total(xyz, q, box, beta, K, 6.0)
    s = system.System(name="pts", box=box.astype(np.float32), crd=xyz_q.astype(np.float32),
                      charge=qq.astype(np.float32), mass=np.ones(8, np.float32),
                      lj_type=np.zeros(8, np.int32), lj_A=np.zeros(1, np.float32),
                      lj_B=np.zeros(1, np.float32), excl_start=np.zeros(8, np.int32),
                      excl_numbers=np.zeros(8, np.int32), excl_list=np.zeros(0, np.int32))
    u = oracle.crd_to_uint(oracle.periodic_map(s.crd, s.box), s.box)
    F = oracle.pme_reciprocal(s, u, K, beta)["frc"]
    for i, d in ((0, 0), (3, 1), (7, 2)):
        up, um, hh = shift_uint(u, i, d, 1e-4, s.box)
        e = [oracle.pme_reciprocal(s, uu, K, beta)["e_rec"] for uu in (up, um)]
        assert F[i, d] == pytest.approx(-(e[0] - e[1]) / hh, rel=1e-6, abs=1e-6)


def test_spme_charge_grid_sums_to_total_charge():
    s = system.config(0)
    u = oracle.crd_to_uint(oracle.periodic_map(s.crd, s.box), s.box)
    r = oracle.pme_reciprocal(s, u, (20, 20, 20), 0.35, want_grids=True)
    assert r["Q"].sum() == pytest.approx(float(s.charge.astype(np.float64).sum()), abs=1e-9)
    q1 = s.charge.astype(np.float64).copy()
    q1[:] = 0
    q1[5] = 1.0
    r1 = oracle.pme_reciprocal(s, u, (20, 20, 20), 0.35, charge=q1.astype(np.float32), want_grids=True)
    assert r1["Q"].sum() == pytest.approx(1.0, abs=1e-12)
    assert (r1["Q"] != 0).sum() <= 64


def test_self_and_lrc_closed_forms():
    q = np.array([1.0, -2.0, 0.5], np.float32)
    box = np.array([10.0, 10.0, 10.0], np.float32)
    es, en = oracle.pme_self(q, 0.3, box)
    assert es == pytest.approx(-0.3 / math.sqrt(math.pi) * 5.25, rel=1e-14)
    assert en == pytest.approx(-math.pi / (2 * 0.09 * 1000) * 0.25, rel=1e-14)
    s = system.config(0)
    f = oracle.lj_lrc_factor(s)
    nO = int((s.lj_type == 0).sum())
    want = -(2.0 / 3.0) * math.pi * nO * nO * float(s.lj_B[0]) / 6.0 / s.cutoff ** 3
    assert f == pytest.approx(want, rel=1e-12)
