"""ctypes bindings to the FP64 oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I, F, D, Lg = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_long
        sig = {
            "orc_crd_to_uint": (None, [I, P, P, P]),
            "orc_periodic_map": (None, [I, P, P]),
            "orc_nl_brute": (Lg, [I, P, P, F, F, P, P, P, P, Lg]),
            "orc_nl_grid": (Lg, [I, P, P, F, F, P, P, P, P, Lg, P]),
            "orc_direct": (None, [I, P, P, P, P, P, P, D, D, Lg, P, P, P, P, P, P]),
            "orc_excluded": (D, [I, P, P, P, D, P, P, P, P, P]),
            "orc_direct_soft": (None, [I, P, P, P, P, P, P, P, P, P, D, D, D, D, D, P, D, D, Lg, P,
                                       P, P, P, P, P]),
            "orc_lj_soft_lrc_factor": (D, [I, P, P, I, I, P, P, D, D]),
            "orc_energy_grad": (None, [I, P, P, P, I, P, P, P, D, D, Lg, P, P, P, P, I, I, I, P, P, P]),
            "orc_loss_grad": (D, [I, P, P, P, I, P, P, P, D, D, Lg, P, P, P, P, I, I, I, P, P, P, P, P]),
            "orc_pme_beta": (D, [D, D]),
            "orc_bspline_mod2": (D, [I, I]),
            "orc_pme_bc_full": (None, [I, I, I, P, D, P]),
            "orc_fft3_r2c": (None, [I, I, I, P, P]),
            "orc_fft3_c2r": (None, [I, I, I, P, P]),
            "orc_pme_reciprocal": (D, [I, P, P, P, I, I, I, D, P, P, P]),
            "orc_pme_self": (None, [I, P, D, P, P]),
            "orc_lj_lrc_factor": (D, [I, P, I, P, D]),
            "orc_ewald_exact": (D, [I, P, P, P, D, D, I, P]),
            "orc_num_threads": (I, []),
            "orc_settle_geometry": (None, [D, D, D, D, D, D, P]),
            "orc_settle_groups": (None, [I, I, P, P, P, P, P, P, P, P, P]),
            "orc_md_iteration": (D, [I, P, P, P, P, P, P, P, D, D, D, D, I, P, P, I, P, P]),
            "orc_fep_windows": (None, [I, P, P, P, P, P, P, P, P, P, P, D, D, D, D, D, P, D, D, Lg, P,
                                       I, P, I, P]),
            "orc_excluded_cross": (D, [I, P, P, P, P, D, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def crd_to_uint(crd, box):
    crd = _c(crd, np.float32)
    box = _c(box, np.float32)
    out = np.zeros(crd.shape, np.uint32)
    lib().orc_crd_to_uint(crd.shape[0], _p(crd), _p(box), _p(out))
    return out


def periodic_map(crd, box):
    crd = _c(crd, np.float32).copy()
    lib().orc_periodic_map(crd.shape[0], _p(crd), _p(_c(box, np.float32)))
    return crd


def _excl(sys_):
    return (_c(sys_.excl_start, np.int32), _c(sys_.excl_numbers, np.int32),
            _c(sys_.excl_list, np.int32) if len(sys_.excl_list) else np.zeros(1, np.int32))


def nl_brute(sys_, uint_crd):
    """Sorted (P,2) int32 array of pairs i<j in the neighbor list."""
    u = _c(uint_crd, np.uint32)
    box = _c(sys_.box, np.float32)
    es, en, el = _excl(sys_)
    m = lib().orc_nl_brute(sys_.n, _p(u), _p(box), sys_.cutoff, sys_.skin,
                           _p(es), _p(en), _p(el), None, 0)
    out = np.zeros((max(m, 1), 2), np.int32)
    lib().orc_nl_brute(sys_.n, _p(u), _p(box), sys_.cutoff, sys_.skin,
                       _p(es), _p(en), _p(el), _p(out), m)
    return out[:m]


def nl_grid(sys_, crd):
    crd = _c(crd, np.float32)
    box = _c(sys_.box, np.float32)
    es, en, el = _excl(sys_)
    dims = np.zeros(3, np.int32)
    m = lib().orc_nl_grid(sys_.n, _p(crd), _p(box), sys_.cutoff, sys_.skin,
                          _p(es), _p(en), _p(el), None, 0, _p(dims))
    out = np.zeros((max(m, 1), 2), np.int32)
    lib().orc_nl_grid(sys_.n, _p(crd), _p(box), sys_.cutoff, sys_.skin,
                      _p(es), _p(en), _p(el), _p(out), m, _p(dims))
    return out[:m], tuple(int(d) for d in dims)


def direct(sys_, uint_crd, pairs, beta, charge=None):
    """FP64 LJ + PME-direct terms. Returns dict of forces, per-atom arrays,
    totals (E_lj, E_direct, W_lj, n_in_cutoff)."""
    n = sys_.n
    u = _c(uint_crd, np.uint32)
    q = _c(sys_.charge if charge is None else charge, np.float32)
    t = _c(sys_.lj_type, np.int32)
    A = _c(sys_.lj_A, np.float32)
    B = _c(sys_.lj_B, np.float32)
    pr = _c(pairs, np.int32)
    frc = np.zeros((n, 3))
    ae = np.zeros(n)
    av = np.zeros(n)
    ad = np.zeros(n)
    tot = np.zeros(4)
    lib().orc_direct(n, _p(u), _p(q), _p(t), _p(A), _p(B), _p(_c(sys_.box, np.float32)),
                     float(sys_.cutoff), float(beta), pr.shape[0], _p(pr), _p(frc),
                     _p(ae), _p(av), _p(ad), _p(tot))
    return dict(frc=frc, atom_energy=ae, atom_virial=av, atom_direct=ad,
                e_lj=tot[0], e_direct=tot[1], w_lj=tot[2], n_in=int(tot[3]))


def direct_soft(fs, uint_crd, pairs, beta, charge=None, lam=None):
    """FP64 soft-core LJ + PME-direct terms of a FepSystem (LJ_soft_core.cu
    restated in orc_direct_soft); same return dict as direct()."""
    n = fs.n
    u = _c(uint_crd, np.uint32)
    q = _c(fs.charge if charge is None else charge, np.float32)
    pr = _c(pairs, np.int32)
    frc = np.zeros((n, 3))
    ae = np.zeros(n)
    av = np.zeros(n)
    ad = np.zeros(n)
    tot = np.zeros(4)
    lib().orc_direct_soft(
        n, _p(u), _p(q), _p(_c(fs.lj_type, np.int32)), _p(_c(fs.lj_type_B, np.int32)),
        _p(_c(fs.mask, np.int32)), _p(_c(fs.lj_A, np.float32)), _p(_c(fs.lj_B, np.float32)),
        _p(_c(fs.lj_A_B, np.float32)), _p(_c(fs.lj_B_B, np.float32)),
        float(fs.lam if lam is None else lam), float(fs.sc_alpha), float(fs.sc_p), float(fs.sc_sigma),
        float(fs.sc_sigma_min), _p(_c(fs.box, np.float32)), float(fs.cutoff), float(beta),
        pr.shape[0], _p(pr), _p(frc), _p(ae), _p(av), _p(ad), _p(tot))
    return dict(frc=frc, atom_energy=ae, atom_virial=av, atom_direct=ad,
                e_lj=tot[0], e_direct=tot[1], w_lj=tot[2], n_in=int(tot[3]))


def lj_soft_lrc_factor(fs, lam=None):
    return lib().orc_lj_soft_lrc_factor(
        fs.n, _p(_c(fs.lj_type, np.int32)), _p(_c(fs.lj_type_B, np.int32)), fs.ntypes, fs.ntypes_B,
        _p(_c(fs.lj_B, np.float32)), _p(_c(fs.lj_B_B, np.float32)),
        float(fs.lam if lam is None else lam), float(fs.cutoff))


def _grad_args(sys_, uint_crd, pairs, beta, K, charge, A, B):
    es, en, el = _excl(sys_)
    return [sys_.n, _p(_c(uint_crd, np.uint32)), _p(_c(sys_.charge if charge is None else charge, np.float32)),
            _p(_c(sys_.lj_type, np.int32)), sys_.ntypes, _p(_c(sys_.lj_A if A is None else A, np.float32)),
            _p(_c(sys_.lj_B if B is None else B, np.float32)), _p(_c(sys_.box, np.float32)), float(sys_.cutoff),
            float(beta), len(pairs), _p(_c(pairs, np.int32)), _p(es), _p(en), _p(el), int(K[0]), int(K[1]),
            int(K[2])], (es, en, el)


def energy_grad(sys_, uint_crd, pairs, beta, K, charge=None, A=None, B=None):
    """FP64 dE/dq (n), dE/dA and dE/dB (stored x12 / x6 triangular tables) of
    the total nonbonded energy (orc_energy_grad)."""
    args, keep = _grad_args(sys_, uint_crd, pairs, beta, K, charge, A, B)
    P = len(sys_.lj_A)
    gq, gA, gB = np.zeros(sys_.n), np.zeros(P), np.zeros(P)
    lib().orc_energy_grad(*args, _p(gq), _p(gA), _p(gB))
    return gq, gA, gB


def loss_grad(sys_, uint_crd, pairs, beta, K, F, T, charge=None, A=None, B=None):
    """Force-matching loss L = sum |F - T|^2 and its FP64 gradients
    (orc_loss_grad); F are the forces of the same parameters."""
    args, keep = _grad_args(sys_, uint_crd, pairs, beta, K, charge, A, B)
    P = len(sys_.lj_A)
    gq, gA, gB = np.zeros(sys_.n), np.zeros(P), np.zeros(P)
    L = lib().orc_loss_grad(*args, _p(_c(F, np.float64)), _p(_c(T, np.float64)), _p(gq), _p(gA), _p(gB))
    return L, gq, gA, gB


def total_energy(sys_, charge=None, A=None, B=None, beta=None, K=None):
    """E_LJ + E_lrc + E_direct + E_correction + E_reciprocal + E_self (FP64)."""
    import dataclasses
    s = sys_
    if A is not None or B is not None:
        s = dataclasses.replace(sys_, lj_A=np.asarray(A if A is not None else sys_.lj_A, np.float32),
                                lj_B=np.asarray(B if B is not None else sys_.lj_B, np.float32))
    r = full_reference(s, charge=charge, beta=beta, K=K)
    return r["e_lj"] + r["e_lrc"] + r["e_direct"] + r["e_corr"] + r["e_rec"] + r["e_self"]


def excluded(sys_, uint_crd, beta, charge=None):
    n = sys_.n
    u = _c(uint_crd, np.uint32)
    q = _c(sys_.charge if charge is None else charge, np.float32)
    es, en, el = _excl(sys_)
    frc = np.zeros((n, 3))
    ac = np.zeros(n)
    e = lib().orc_excluded(n, _p(u), _p(q), _p(_c(sys_.box, np.float32)), float(beta),
                           _p(es), _p(en), _p(el), _p(frc), _p(ac))
    return dict(frc=frc, atom_corr=ac, e_corr=e)


def pme_beta(rc, tol=1e-5):
    return lib().orc_pme_beta(float(rc), float(tol))


def bspline_mod2(k, K):
    return lib().orc_bspline_mod2(int(k), int(K))


def pme_bc_full(K, box, beta):
    bc = np.zeros(tuple(K))
    lib().orc_pme_bc_full(K[0], K[1], K[2], _p(_c(box, np.float32)), float(beta), _p(bc))
    return bc


def fft3_r2c(a):
    a = _c(a, np.float64)
    Kx, Ky, Kz = a.shape
    out = np.zeros((Kx, Ky, Kz // 2 + 1, 2))
    lib().orc_fft3_r2c(Kx, Ky, Kz, _p(a), _p(out))
    return out[..., 0] + 1j * out[..., 1]


def fft3_c2r(c, Kz):
    Kx, Ky, hz = c.shape
    buf = _c(np.stack([c.real, c.imag], -1), np.float64)
    out = np.zeros((Kx, Ky, Kz))
    lib().orc_fft3_c2r(Kx, Ky, Kz, _p(buf), _p(out))
    return out


FEP_FIELDS = ["vdw_intersys", "vdw_intrasys", "coul_direct_intersys", "coul_direct_intrasys",
              "dhdl_lj", "dhdl_direct", "dhdl_direct_exact"]


def fep_windows(fs, uint_crd, pairs, beta, lambdas, charge=None, dq=None, lambda0=None, truncate_soft=True):
    """FP64 FEP / TI pair terms of a FepSystem at every lambda window
    (orc_fep_windows): array (nwin, 7) in FEP_FIELDS order. Window charges
    are charge + (lambda - lambda0) dq (charge: fs.charge, lambda0: fs.lam).
    truncate_soft=False drops the energy kernel's soft-core-radius cutoff."""
    u = _c(uint_crd, np.uint32)
    q = _c(fs.charge if charge is None else charge, np.float32)
    lams = _c(np.atleast_1d(lambdas), np.float64)
    out = np.zeros((len(lams), 7))
    lib().orc_fep_windows(
        fs.n, _p(u), _p(q), _p(None if dq is None else _c(dq, np.float32)), _p(_c(fs.lj_type, np.int32)),
        _p(_c(fs.lj_type_B, np.int32)), _p(_c(fs.mask, np.int32)), _p(_c(fs.lj_A, np.float32)),
        _p(_c(fs.lj_B, np.float32)), _p(_c(fs.lj_A_B, np.float32)), _p(_c(fs.lj_B_B, np.float32)),
        float(fs.lam if lambda0 is None else lambda0), float(fs.sc_alpha), float(fs.sc_p), float(fs.sc_sigma),
        float(fs.sc_sigma_min), _p(_c(fs.box, np.float32)), float(fs.cutoff), float(beta),
        _c(pairs, np.int32).shape[0], _p(_c(pairs, np.int32)), len(lams), _p(lams), int(bool(truncate_soft)),
        _p(out))
    return out


def excluded_cross(sys_, uint_crd, beta, charge, dq):
    """-sum over excluded pairs (q_i dq_j + dq_i q_j) erf(b r)/r, FP64."""
    es, en, el = _excl(sys_)
    return lib().orc_excluded_cross(sys_.n, _p(_c(uint_crd, np.uint32)), _p(_c(charge, np.float32)),
                                    _p(_c(dq, np.float32)), _p(_c(sys_.box, np.float32)), float(beta),
                                    _p(es), _p(en), _p(el))


def pme_cross_reciprocal(sys_, uint_crd, K, beta, charge, dq):
    """The reference's cross reciprocal term (TI_core.cu:686-697): the grid of
    dq spread like the charges, dotted with the convolved grid of the
    charges, FP64."""
    a = pme_reciprocal(sys_, uint_crd, K, beta, charge, want_grids=True)
    b = pme_reciprocal(sys_, uint_crd, K, beta, dq, want_grids=True)
    return float((b["Q"] * a["conv"]).sum())


def pme_reciprocal(sys_, uint_crd, K, beta, charge=None, want_grids=False):
    n = sys_.n
    u = _c(uint_crd, np.uint32)
    q = _c(sys_.charge if charge is None else charge, np.float32)
    frc = np.zeros((n, 3))
    Q = np.zeros(tuple(K)) if want_grids else None
    C = np.zeros(tuple(K)) if want_grids else None
    e = lib().orc_pme_reciprocal(n, _p(u), _p(q), _p(_c(sys_.box, np.float32)), K[0], K[1], K[2],
                                 float(beta), _p(frc), _p(Q), _p(C))
    out = dict(frc=frc, e_rec=e)
    if want_grids:
        out.update(Q=Q, conv=C)
    return out


def pme_self(charge, beta, box):
    q = _c(charge, np.float32)
    out = np.zeros(2)
    lib().orc_pme_self(len(q), _p(q), float(beta), _p(_c(box, np.float32)), _p(out))
    return out[0], out[1]


def lj_lrc_factor(sys_):
    return lib().orc_lj_lrc_factor(sys_.n, _p(_c(sys_.lj_type, np.int32)), sys_.ntypes,
                                   _p(_c(sys_.lj_B, np.float32)), float(sys_.cutoff))


def ewald_exact(xyz, q, box, beta, rreal, kmax):
    xyz = _c(xyz, np.float64)
    q = _c(q, np.float64)
    frc = np.zeros(xyz.shape)
    e = lib().orc_ewald_exact(len(q), _p(xyz), _p(q), _p(_c(box, np.float32)), float(beta),
                              float(rreal), int(kmax), _p(frc))
    return e, frc


def num_threads():
    return lib().orc_num_threads()


def full_reference(sys_, crd=None, charge=None, beta=None, K=None):
    """Everything the force path produces for one configuration, FP64.

    NL from brute force on the wrapped coordinates, then direct, exclusion,
    reciprocal, self/neutralising and long-range terms."""
    crd = sys_.crd if crd is None else crd
    crd = periodic_map(crd, sys_.box)
    u = crd_to_uint(crd, sys_.box)
    if beta is None:
        beta = float(np.float32(pme_beta(sys_.cutoff, sys_.pme_tolerance)))
    if K is None:
        from sponge_amd.system import fft_size
        K = tuple(sys_.fft[d] if sys_.fft[d] > 0 else fft_size(sys_.box[d]) for d in range(3))
    pairs = nl_brute(sys_, u)
    q = sys_.charge if charge is None else charge
    soft = getattr(sys_, "lj_type_B", None) is not None
    d = direct_soft(sys_, u, pairs, beta, q) if soft else direct(sys_, u, pairs, beta, q)
    x = excluded(sys_, u, beta, q)
    r = pme_reciprocal(sys_, u, K, beta, q)
    es, en = pme_self(q, beta, sys_.box)
    V = float(np.prod(sys_.box.astype(np.float64)))
    lrc = (lj_soft_lrc_factor(sys_) if soft else lj_lrc_factor(sys_)) / V
    frc = d["frc"] + x["frc"] + r["frc"]
    return dict(uint=u, crd=crd, pairs=pairs, beta=beta, K=K, frc=frc,
                frc_direct=d["frc"], frc_excl=x["frc"], frc_rec=r["frc"],
                e_lj=d["e_lj"], e_direct=d["e_direct"], w_lj=d["w_lj"],
                e_corr=x["e_corr"], e_rec=r["e_rec"], e_self=es + en,
                e_lrc=lrc, atom_energy=d["atom_energy"], atom_virial=d["atom_virial"],
                atom_direct=d["atom_direct"], atom_corr=x["atom_corr"], n_in=d["n_in"])


# ---------------------------------------------------------------------------
# Integrator and SETTLE (orc_settle_* / orc_md_iteration).

def settle_geometry(rAB, rAC, rBC, mA, mB, mC):
    g = np.zeros(5)
    lib().orc_settle_geometry(float(rAB), float(rAC), float(rBC), float(mA), float(mB), float(mC), _p(g))
    return g


def settle_groups(n, pair_i, pair_j, pair_r, mass):
    """SETTLE's rigid triangles / pairs found in a constraint-pair list."""
    pi, pj = _c(pair_i, np.int32), _c(pair_j, np.int32)
    pr, m = _c(pair_r, np.float32), _c(mass, np.float32)
    k = len(pi)
    tri = np.zeros((max(k, 1), 3), np.int32)
    geo = np.zeros((max(k, 1), 5))
    prs = np.zeros((max(k, 1), 2), np.int32)
    plen = np.zeros(max(k, 1))
    cnt = np.zeros(3, np.int32)
    lib().orc_settle_groups(int(n), k, _p(pi), _p(pj), _p(pr), _p(m), _p(tri), _p(geo), _p(prs), _p(plen),
                            _p(cnt))
    nt, npr = int(cnt[0]), int(cnt[1])
    return dict(tri=tri[:nt].copy(), geo=geo[:nt].copy(), pairs=prs[:npr].copy(), plen=plen[:npr].copy(),
                rest=int(cnt[2]))


def md_iteration(crd, vel, frc, mass, uint_crd, box, dt, groups, vmax=0.0, vfac=1.0, xfac=1.0):
    """One leap-frog + SETTLE step in FP64; crd / vel (float64, (N,3)) are
    updated in place. Returns the constraint virial sum."""
    assert crd.dtype == np.float64 and vel.dtype == np.float64
    assert crd.flags.c_contiguous and vel.flags.c_contiguous
    m = _c(mass, np.float64)
    im = np.where(m > 0, 1.0 / np.where(m > 0, m, 1.0), 0.0)
    f = _c(frc, np.float64)
    u = _c(uint_crd, np.uint32)
    tri, geo = _c(groups["tri"], np.int32), _c(groups["geo"], np.float64)
    prs, plen = _c(groups["pairs"], np.int32), _c(groups["plen"], np.float64)
    return lib().orc_md_iteration(len(m), _p(crd), _p(vel), _p(f), _p(im), _p(m), _p(u), _p(_c(box, np.float32)),
                                  float(dt), float(vmax), float(vfac), float(xfac), len(tri), _p(tri), _p(geo),
                                  len(prs), _p(prs), _p(plen))


def md_run(sys_, vel0, nsteps, dt, groups, crd0=None, beta=None, K=None, record=None):
    """NVE trajectory of main.cu's loop (force, then leap-frog + SETTLE) in
    FP64 state, forces from full_reference at the FP32-rounded coordinates.
    Returns (crd, vel, log) with per-step potential and kinetic energies."""
    crd = np.array(sys_.crd if crd0 is None else crd0, np.float64)
    vel = np.array(vel0, np.float64)
    m = sys_.mass.astype(np.float64)
    log = []
    for step in range(nsteps):
        ref = full_reference(sys_, crd=crd.astype(np.float32), beta=beta, K=K)
        beta, K = ref["beta"], ref["K"]
        pot = ref["e_lj"] + ref["e_lrc"] + ref["e_direct"] + ref["e_corr"] + ref["e_rec"] + ref["e_self"]
        ke_before = 0.5 * float((m[:, None] * vel * vel).sum())
        md_iteration(crd, vel, ref["frc"], m, ref["uint"], sys_.box, dt, groups)
        ke_after = 0.5 * float((m[:, None] * vel * vel).sum())
        log.append(dict(potential=pot, ke_half_before=ke_before, ke_half_after=ke_after))
        if record is not None:
            record(step, crd, vel)
    return crd, vel, log


# ---------------------------------------------------------------------------
# Pressure and barostat coupling.

def virial(ref):
    """The virial the reference accumulates for the pressure of one force step
    (Lennard_Jones_force.cu:1178-1191, PME_force.cu:753-771, MD_core.cu:
    1777-1795): sum(-f_LJ r^2) + 6 E_lrc + E_direct + E_correction + E_self +
    E_reciprocal, from a full_reference dict."""
    return (ref["w_lj"] + 6.0 * ref["e_lrc"] + ref["e_direct"] + ref["e_corr"] + ref["e_self"]
            + ref["e_rec"])


def pressure(vel, mass, W, box):
    """Calculate_Pressure_Cuda (MD_core.cu:183-186) with MD_Atom_Ek (:241-250),
    FP64: (2 Ek + W) / (3 V)."""
    v = np.asarray(vel, np.float64)
    ek = 0.5 * float((np.asarray(mass, np.float64)[:, None] * v * v).sum())
    V = float(np.prod(np.asarray(box, np.float64)))
    return (2.0 * ek + W) / (3.0 * V), ek


def barostat_run(sys_, vel0, nsteps, dt_ps, groups, kind="berendsen", target_pressure_bar=1.0, tau=1.0,
                 compressibility=4.5e-5, update_interval=10, dV_dt=0.0, beta=None, K=None):
    """main.cu's NVE loop with a barostat, FP64 state. The pressure of a force
    step plus SETTLE's constraint part (0.33333 W_c / V, SETTLE.cu:507-525)
    drives, after the integration:
    berendsen (main.cu:334-354, barostats/Berendsen_barostat.cu) at steps that
      are multiples of update_interval (counted from 1):
      mu = 1 - interval * compressibility * dt / tau / 3 * (p0 - p), dt and
      tau in ps; coordinates scaled by mu;
    andersen (main.cu:356-372, barostats/andersen_barostat.cu) every step:
      dV/dt += (p - p0) V0 compressibility / tau^2 (tau in internal time),
      V' = V + dV/dt dt, mu = cbrt(V' / V); coordinates scaled by mu,
      velocities by 1 / mu;
    then the box is scaled by mu. Returns (crd, vel, box, log)."""
    import dataclasses
    p0 = target_pressure_bar * 0.00001439506089041446
    comp = compressibility * 6.946827162543585e4
    crd = np.array(sys_.crd, np.float64)
    vel = np.array(vel0, np.float64)
    box = np.array(sys_.box, np.float64)
    m = sys_.mass.astype(np.float64)
    dt = dt_ps * 20.455
    V0 = float(np.prod(box))
    inv_w = V0 * comp / (tau * 20.455) ** 2
    interval = 1 if kind == "andersen" else update_interval
    log = []
    for step in range(1, nsteps + 1):
        s = dataclasses.replace(sys_, box=box.astype(np.float32))
        ref = full_reference(s, crd=crd.astype(np.float32), beta=beta, K=K)
        beta, K = ref["beta"], ref["K"]
        need = step % interval == 0
        if need:
            p, ek = pressure(vel, m, virial(ref), s.box)
        wc = md_iteration(crd, vel, ref["frc"], m, ref["uint"], s.box, dt, groups)
        if need:
            V = float(np.prod(s.box.astype(np.float64)))
            p += 0.33333 * wc / V
            if kind == "andersen":
                dV_dt += (p - p0) * inv_w
                mu = np.cbrt((V + dV_dt * dt) / V)
                vel /= mu
            else:
                mu = 1.0 - update_interval * comp * dt_ps / tau / 3.0 * (p0 - p)
            crd *= mu
            box *= mu
            log.append(dict(step=step, pressure=p, mu=mu))
    return crd, vel, box, log


# ---------------------------------------------------------------------------
# OpenMP FP32 CPU baseline of the MD step (cpu_md.cpp): bench.py's
# cpu_baseline leg and its tests.

_md_lib = None


def _cpu_md_lib():
    global _md_lib
    if _md_lib is None:
        path = os.path.join(_HERE, "_build", "libcpumd.so")
        if not os.path.exists(path):
            build(force=True)
        L = ctypes.CDLL(path)
        P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        L.cpu_md_create.restype = P
        L.cpu_md_create.argtypes = [I, P, P, P, P, P, P, I, P, P, P, P, P, F, F, I, I, I, I, F, P, F, I, F, F]
        L.cpu_md_destroy.argtypes = [P]
        L.cpu_md_forces.argtypes = [P, P]
        L.cpu_md_run.restype = I
        L.cpu_md_run.argtypes = [P, I]
        L.cpu_md_state.argtypes = [P, P, P]
        L.cpu_md_threads.restype = I
        _md_lib = L
    return _md_lib


class CpuMd:
    """The whole NVE step (list every `refresh` steps, LJ + PME direct,
    exclusion correction, SPME reciprocal with its own FFT, leap-frog +
    SETTLE) in FP32 with OpenMP on the host cores."""

    def __init__(self, sys_, vel, dt, refresh=20, beta=None, K=None):
        from sponge_amd.system import fft_size, TIP3P_ROH, TIP3P_HOH_DEG
        import math
        if beta is None:
            beta = float(np.float32(pme_beta(sys_.cutoff, sys_.pme_tolerance)))
        if K is None:
            K = tuple(sys_.fft[d] if sys_.fft[d] > 0 else fft_size(sys_.box[d]) for d in range(3))
        self.K, self.beta, self.n = tuple(int(k) for k in K), float(beta), sys_.n
        bc = pme_bc_full(self.K, sys_.box, beta)
        es, en, el = _excl(sys_)
        self._keep = [_c(a, dt_) for a, dt_ in (
            (sys_.box, np.float32), (sys_.crd, np.float32), (vel, np.float32), (sys_.mass, np.float32),
            (sys_.charge, np.float32), (sys_.lj_type, np.int32), (sys_.lj_A, np.float32),
            (sys_.lj_B, np.float32))] + [es, en, el, _c(bc, np.float64)]
        k = self._keep
        r_hh = 2.0 * TIP3P_ROH * math.sin(math.radians(TIP3P_HOH_DEG) / 2)
        self.h = _cpu_md_lib().cpu_md_create(
            sys_.n, _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), _p(k[4]), _p(k[5]), sys_.ntypes, _p(k[6]), _p(k[7]),
            _p(k[8]), _p(k[9]), _p(k[10]), sys_.cutoff, sys_.skin, int(refresh), *self.K, self.beta, _p(k[11]),
            float(dt), int(sys_.waters), TIP3P_ROH, r_hh)

    def forces(self):
        out = np.zeros((self.n, 3), np.float32)
        _cpu_md_lib().cpu_md_forces(self.h, _p(out))
        return out

    def run(self, steps):
        """`steps` NVE steps; raises if a neighbor list overflowed."""
        if _cpu_md_lib().cpu_md_run(self.h, int(steps)):
            raise RuntimeError("CPU baseline: neighbor list capacity (1024 per atom) exceeded")

    def state(self):
        c = np.zeros((self.n, 3), np.float32)
        v = np.zeros((self.n, 3), np.float32)
        _cpu_md_lib().cpu_md_state(self.h, _p(c), _p(v))
        return c, v

    @staticmethod
    def threads():
        return _cpu_md_lib().cpu_md_threads()

    def __del__(self):
        if getattr(self, "h", None):
            _cpu_md_lib().cpu_md_destroy(self.h)
            self.h = None
