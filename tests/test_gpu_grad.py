"""The HIP adjoint (sponge_nb_energy_gradient / _force_matching_gradient)
against the FP64 oracle's analytic gradients (pinned by finite differences in
tests/test_oracle_grad.py) and, on the 23,570-atom water box with ions, against
FP64 central differences of the oracle's energy and loss on 64 random
parameters (h = 1e-4 relative). Target forces come from the same system with
charges perturbed by N(0, 0.02) e. Bar: 1e-4 relative.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from sponge_amd import system

pytestmark = pytest.mark.gpu

TOL = 1e-4


def small(name):
    if name == "water648":
        return system.config(0)
    return system.small_random(300, (14.0, 15.0, 16.0), seed=11, cutoff=5.0, skin=1.0)


def make(sys_):
    import torch
    from sponge_amd import nb
    eng = nb.NonbondedEngine(sys_, refresh_interval=0)
    eng.update_neighbor_list(forced=1)
    eng.crd_to_uint()
    torch.cuda.synchronize()
    return eng


def gpu_energy_grad(eng):
    import torch
    P = len(eng.sys.lj_A)
    gq = torch.zeros(eng.n, dtype=torch.float64, device="cuda")
    gA = torch.zeros(P, dtype=torch.float64, device="cuda")
    gB = torch.zeros(P, dtype=torch.float64, device="cuda")
    eng.energy_gradient(gq, gA, gB)
    torch.cuda.synchronize()
    return gq.cpu().numpy(), gA.cpu().numpy(), gB.cpu().numpy()


def gpu_forces(eng):
    import torch
    eng.force_step(0, 0, from_crd=True, overwrite=True)
    torch.cuda.synchronize()
    return eng.frc.clone()


def gpu_loss_grad(eng, target, forces=None):
    """forces=None: the GPU's own forces of the current parameters; else the
    given forces (FP32) are the F of the loss."""
    import torch
    P = len(eng.sys.lj_A)
    frc = gpu_forces(eng)
    if forces is not None:
        frc = torch.as_tensor(forces, dtype=torch.float32, device="cuda").contiguous()
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    gq = torch.zeros(eng.n, dtype=torch.float64, device="cuda")
    gA = torch.zeros(P, dtype=torch.float64, device="cuda")
    gB = torch.zeros(P, dtype=torch.float64, device="cuda")
    tgt = torch.as_tensor(target, dtype=torch.float32, device="cuda").contiguous()
    eng.force_matching_gradient(frc, tgt, loss, gq, gA, gB)
    torch.cuda.synchronize()
    return frc.cpu().numpy().astype(np.float64), float(loss.item()), gq.cpu().numpy(), gA.cpu().numpy(), \
        gB.cpu().numpy()


def ref_setup(sys_, eng):
    crd = oracle.periodic_map(sys_.crd.copy(), sys_.box)
    u = oracle.crd_to_uint(crd, sys_.box)
    pairs = oracle.nl_brute(sys_, u)
    K = (eng.pme.fftx, eng.pme.ffty, eng.pme.fftz)
    return u, pairs, float(eng.pme.beta), K


def check_vec(g, ref, what):
    scale = float(np.sqrt((ref ** 2).mean())) + 1e-30
    err = np.abs(g - ref) / np.maximum(np.abs(ref), scale)
    assert err.max() < TOL, (what, float(err.max()), int(err.argmax()))


def target_forces(sys_, beta, K, seed=5):
    rng = np.random.default_rng(seed)
    qt = (sys_.charge + rng.normal(0.0, 0.02 * system.CHARGE_SCALE, sys_.n)).astype(np.float32)
    # the GPU takes FP32 targets: both sides use the same rounded values (near
    # pairs of the random systems make F - T a small difference of large forces)
    return oracle.full_reference(sys_, charge=qt, beta=beta, K=K)["frc"].astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("name", ["water648", "random300"])
def test_energy_gradient_matches_oracle(name):
    sys_ = small(name)
    print(f"[grad] energy {name}", flush=True)
    eng = make(sys_)
    u, pairs, beta, K = ref_setup(sys_, eng)
    gq, gA, gB = gpu_energy_grad(eng)
    rq, rA, rB = oracle.energy_grad(sys_, u, pairs, beta, K)
    check_vec(gq, rq, "charge")
    check_vec(gA, rA, "A")
    check_vec(gB, rB, "B")


@pytest.mark.parametrize("name", ["water648", "random300"])
def test_loss_gradient_matches_oracle(name):
    sys_ = small(name)
    if name.startswith("random"):
        sys_ = dataclasses.replace(sys_, lj_A=(sys_.lj_A * 1e-8).astype(np.float32),
                                   lj_B=(sys_.lj_B * 1e-4).astype(np.float32))
    print(f"[grad] loss {name}", flush=True)
    eng = make(sys_)
    u, pairs, beta, K = ref_setup(sys_, eng)
    T = target_forces(sys_, beta, K)
    F, L, gq, gA, gB = gpu_loss_grad(eng, T)
    rL, rq, rA, rB = oracle.loss_grad(sys_, u, pairs, beta, K, F, T)
    assert abs(L - rL) <= 1e-6 * rL
    check_vec(gq, rq, "charge")
    check_vec(gA, rA, "A")
    check_vec(gB, rB, "B")


def _energy(sys_, u, pairs, beta, K, q=None, A=None, B=None):
    s = dataclasses.replace(sys_, lj_A=np.asarray(A if A is not None else sys_.lj_A, np.float32),
                            lj_B=np.asarray(B if B is not None else sys_.lj_B, np.float32))
    q = s.charge if q is None else q
    d = oracle.direct(s, u, pairs, beta, q)
    x = oracle.excluded(s, u, beta, q)
    r = oracle.pme_reciprocal(s, u, K, beta, q)
    es, en = oracle.pme_self(q, beta, s.box)
    V = float(np.prod(s.box.astype(np.float64)))
    return (d["e_lj"] + oracle.lj_lrc_factor(s) / V + d["e_direct"] + x["e_corr"] + r["e_rec"] + es + en,
            d["frc"] + x["frc"] + r["frc"])


def _params(sys_, rng, nq):
    qs = [("q", int(k)) for k in rng.choice(sys_.n, nq, replace=False)]
    tabs = []
    for p in range(len(sys_.lj_A)):
        if sys_.lj_A[p] != 0:
            tabs += [("A", p), ("B", p)]
    return qs + tabs


def _perturbed(sys_, kind, k, sign):
    q, A, B = sys_.charge.copy(), sys_.lj_A.astype(np.float64), sys_.lj_B.astype(np.float64)
    arr = {"q": q, "A": A, "B": B}[kind]
    h = 1e-4 * abs(float(arr[k]))
    arr[k] = arr[k] + sign * h
    return q.astype(np.float32), A.astype(np.float32), B.astype(np.float32)


def test_config5_gradients_vs_central_differences():
    """23,570-atom box: GPU energy and loss gradients against the full FP64
    oracle gradient and against FP64 central differences on 64 parameters
    (52 charges and the A / B entries of the O, Na, Cl pair types)."""
    sys_ = system.config(1)
    print("[grad] water23570", flush=True)
    eng = make(sys_)
    u, pairs, beta, K = ref_setup(sys_, eng)
    gq, gA, gB = gpu_energy_grad(eng)
    rq, rA, rB = oracle.energy_grad(sys_, u, pairs, beta, K)
    check_vec(gq, rq, "charge (energy)")
    check_vec(gA, rA, "A (energy)")
    check_vec(gB, rB, "B (energy)")
    T = target_forces(sys_, beta, K)
    F, L, lq, lA, lB = gpu_loss_grad(eng, T)
    sq, sA, sB = oracle.loss_grad(sys_, u, pairs, beta, K, F, T)[1:]
    check_vec(lq, sq, "charge (loss)")
    check_vec(lA, sA, "A (loss)")
    check_vec(lB, sB, "B (loss)")
    # Central differences below differentiate the loss of the FP64 oracle's
    # forces; w = 2 (F - T) amplifies the FP32 forward error of F by |F|/|F-T|
    # (about 50 here), so the adjoint is given the oracle's forces (rounded to
    # FP32) as its F, isolating the adjoint from the forward rounding.
    F0 = _energy(sys_, u, pairs, beta, K)[1].astype(np.float32)
    _, _, lq, lA, lB = gpu_loss_grad(eng, T, forces=F0)
    rng = np.random.default_rng(7)
    params = _params(sys_, rng, 64 - 12)
    assert len(params) == 64
    gpu_e = {"q": gq, "A": gA, "B": gB}
    gpu_l = {"q": lq, "A": lA, "B": lB}
    for kind, k in params:
        ep, fp = _energy(sys_, u, pairs, beta, K, *_perturbed(sys_, kind, k, +1))
        em, fm = _energy(sys_, u, pairs, beta, K, *_perturbed(sys_, kind, k, -1))
        xp = _perturbed(sys_, kind, k, +1)[("q", "A", "B").index(kind)][k]
        xm = _perturbed(sys_, kind, k, -1)[("q", "A", "B").index(kind)][k]
        dx = float(xp) - float(xm)
        fd_e = (ep - em) / dx
        fd_l = (((fp - T) ** 2).sum() - ((fm - T) ** 2).sum()) / dx  # FP64 forces on both sides
        ge, gl = gpu_e[kind][k], gpu_l[kind][k]
        assert abs(ge - fd_e) <= TOL * max(abs(fd_e), 1e-3 * np.sqrt((gpu_e[kind] ** 2).mean())), (kind, k, ge, fd_e)
        assert abs(gl - fd_l) <= TOL * max(abs(fd_l), 1e-3 * np.sqrt((gpu_l[kind] ** 2).mean())), (kind, k, gl, fd_l)


def test_autograd_functions():
    import torch
    from sponge_amd import grad
    sys_ = system.config(0)
    eng = make(sys_)
    q = torch.as_tensor(sys_.charge, device="cuda").clone().requires_grad_(True)
    A = torch.as_tensor(sys_.lj_A, dtype=torch.float64).clone().requires_grad_(True)
    B = torch.as_tensor(sys_.lj_B, dtype=torch.float64).clone().requires_grad_(True)
    E = grad.nonbonded_energy(eng, q, A, B)
    E.backward()
    gq, gA, gB = gpu_energy_grad(eng)
    np.testing.assert_allclose(q.grad.cpu().numpy(), gq.astype(np.float32), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(A.grad.numpy(), gA, rtol=1e-9, atol=0)
    u, pairs, beta, K = ref_setup(sys_, eng)
    ref = oracle.full_reference(sys_, beta=beta, K=K)
    e_ref = ref["e_lj"] + ref["e_lrc"] + ref["e_direct"] + ref["e_corr"] + ref["e_rec"] + ref["e_self"]
    assert abs(E.item() - e_ref) <= 1e-6 * abs(e_ref)
    T = torch.as_tensor(target_forces(sys_, beta, K), dtype=torch.float32, device="cuda")
    for t in (q, A, B):
        t.grad = None
    L = grad.force_matching_loss(eng, q, A, B, T)
    L.backward()
    F, L2, lq, lA, lB = gpu_loss_grad(eng, T.cpu().numpy())
    assert abs(L.item() - L2) <= 1e-7 * L2  # the kernel forms F - T in FP32
    np.testing.assert_allclose(q.grad.cpu().numpy(), lq.astype(np.float32), rtol=1e-5, atol=1e-6)
    pq = grad.project_net_charge(q.grad)
    assert abs(float(pq.sum())) < 1e-3 * float(pq.abs().sum())
