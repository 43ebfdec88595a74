"""Neighbor-list paths beyond a plain forced rebuild, through the C ABI:

* the skin-triggered rebuild (Neighbor_List_Update without a forced update or
  a refresh interval, neighbor_list.cu:336-351, 405-451, 543-545): decided on
  the device, eagerly and inside a captured hipGraph, with the reference's
  skin_permit scaling of the threshold (neighbor_list.cu:608-610);
* list symmetry: every pair stored from both of its atoms;
* the column-overflow error path (a cell column with more atoms than it can
  hold; the reference's buckets overflow silently, neighbor_list.cu:200-214):
  reported, and the atoms left out get a defined state;
* a PME grid axis above 256 points.
"""
import dataclasses
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


def oracle_pairs(sys_, crd):
    crd = oracle.periodic_map(crd, sys_.box)
    return oracle.nl_brute(sys_, oracle.crd_to_uint(crd, sys_.box))


def _skin_engine(sys_, skin_permit=None):
    from sponge_amd import nb
    eng = nb.NonbondedEngine(sys_, refresh_interval=0)
    if skin_permit is not None:
        eng.nl.Clear()
        eng.nl = nb.NEIGHBOR_LIST(sys_.n, eng.box, sys_.cutoff, sys_.skin, refresh_interval=0,
                                  skin_permit=skin_permit)
    return eng


def _move(eng, atom, dist):
    """Moves one atom by `dist` A along a fixed unit vector (wrapped into
    the box), device-side."""
    import torch
    u = np.array([0.48, -0.6, 0.64])
    u /= np.linalg.norm(u)
    c = eng.crd.cpu().numpy().astype(np.float64)
    c[atom] = np.mod(c[atom] + dist * u, np.asarray(eng.box, np.float64))
    c = c.astype(np.float32)
    c[c >= np.asarray(eng.box, np.float32)] = 0.0
    eng.crd.copy_(torch.from_numpy(c))


@pytest.mark.parametrize("skin_permit", [None, 1.0])
def test_skin_triggered_rebuild_eager(skin_permit):
    """No rebuild while every atom stays within skin_permit' * skin / 2 of its
    position at the last rebuild (skin_permit' = 2 x the mdin value: 1 by
    default, 2 for mdin 1.0); one rebuild, with the oracle's exact set at the
    new positions, once one atom moves past it."""
    import torch
    sys_ = system.config(0)
    eng = _skin_engine(sys_, skin_permit)
    reach = (1.0 if skin_permit is None else 2.0 * skin_permit) * sys_.skin / 2.0
    say(f"[case] skin-triggered, n={sys_.n}, skin_permit={skin_permit}, threshold {reach} A")
    eng.update_neighbor_list(forced=1)
    r0 = eng.nl.status()["rebuilds"]
    atom = 100
    _move(eng, atom, 0.97 * reach)
    eng.update_neighbor_list()
    torch.cuda.synchronize()
    assert eng.nl.status()["rebuilds"] == r0, "rebuilt below the skin threshold"
    _move(eng, atom, 0.06 * reach)   # now 1.03 x the threshold from the last rebuild
    eng.update_neighbor_list()
    torch.cuda.synchronize()
    st = eng.nl.check()
    assert st["rebuilds"] == r0 + 1, "no rebuild past the skin threshold"
    got = eng.nl.pairs()
    ref = oracle_pairs(sys_, eng.crd.cpu().numpy())
    np.testing.assert_array_equal(got, ref)
    # the new positions are the reference for the next check
    eng.update_neighbor_list()
    torch.cuda.synchronize()
    assert eng.nl.status()["rebuilds"] == r0 + 1


def test_skin_triggered_rebuild_in_graph():
    """The skin check and the rebuild it triggers captured into one hipGraph
    (the device decides; no host round trip): replays with unmoved atoms
    rebuild nothing, a replay after an atom moved past the threshold rebuilds
    once and yields the oracle's set; the force step captured after it sees
    the new list."""
    import torch
    sys_ = system.config(0)
    eng = _skin_engine(sys_)
    say(f"[case] skin-triggered in a graph, n={sys_.n}")
    eng.update_neighbor_list(forced=1)
    eng.force_step(0, 0, from_crd=True, overwrite=True)
    torch.cuda.synchronize()
    r0 = eng.nl.status()["rebuilds"]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cur = torch.cuda.current_stream()
        eng.update_neighbor_list(stream=cur)
        eng.force_step(0, 0, from_crd=True, overwrite=True, stream=cur)
    for _ in range(2):
        g.replay()
    torch.cuda.synchronize()
    assert eng.nl.status()["rebuilds"] == r0
    _move(eng, 200, 1.2 * sys_.skin / 2.0)
    g.replay()
    torch.cuda.synchronize()
    st = eng.nl.check()
    assert st["rebuilds"] == r0 + 1
    crd = eng.crd.cpu().numpy()
    np.testing.assert_array_equal(eng.nl.pairs(), oracle_pairs(sys_, crd))
    s2 = dataclasses.replace(sys_, crd=np.ascontiguousarray(crd, np.float32))
    ref = oracle.full_reference(s2, beta=eng.pme.beta, K=(eng.pme.fftx, eng.pme.ffty, eng.pme.fftz))
    e = rel_rms(eng.frc.cpu().numpy(), ref["frc"])
    say(f"    after the triggered rebuild: force rel rms {e:.3e}")
    assert e < 1e-5
    g.replay()
    torch.cuda.synchronize()
    assert eng.nl.status()["rebuilds"] == r0 + 1


@pytest.mark.parametrize("which", [0, 1])
def test_list_is_symmetric(which):
    """Every listed pair is stored from both of its atoms (the pair kernel
    writes only its own atoms' forces): sponge_nl_get_pairs compares the two
    sides and fails with SPONGE_NB_EINTERNAL otherwise; the directed count
    the rebuild kept is twice the unique count."""
    from sponge_amd import nb
    sys_ = system.config(which)
    say(f"[case] symmetry, n={sys_.n}")
    eng = nb.NonbondedEngine(sys_, refresh_interval=0)
    eng.update_neighbor_list(forced=1)
    pairs = eng.nl.pairs()       # raises on an asymmetric list
    st = eng.nl.check()
    assert st["pairs"] == len(pairs)


def _droplet():
    base = system.config(0)
    return dataclasses.replace(base, name="droplet2x",
                               box=(base.box.astype(np.float64) * 2.0).astype(np.float32))


def test_column_overflow_is_reported_and_defined():
    """The 648-atom water box as a droplet in a box twice as long: a cell
    column of the list holds 128 atoms and the droplet puts 130 in one. The
    rebuild reports SPONGE_NL_OVERFLOW_COLUMN (status, check() raises, the
    pair export fails with the capacity code); the atoms left out take no
    part in the step (no pairs, zero charge on the PME grid, zero force when
    forces are assigned, also over repeated steps), and every other atom gets
    finite forces equal to the oracle's for the system without them."""
    import torch
    from sponge_amd import nb
    sys_ = _droplet()
    say(f"[case] column overflow droplet n={sys_.n} box={sys_.box.tolist()}")
    eng = nb.NonbondedEngine(sys_, refresh_interval=20)
    eng.nl.set_prune(0)
    eng.update_neighbor_list(forced=1)
    st = eng.nl.status()
    say(f"    status {st}")
    assert st["overflow"] & nb.OVERFLOW_COLUMN
    assert not st["ok"]
    with pytest.raises(nb.SpongeError):
        eng.nl.check()
    with pytest.raises(nb.SpongeError):
        eng.nl.pairs()
    # the full list every step, then the per-slot inner lists (pruned every
    # 4 steps to rc + skin / 2)
    for prune in (0, 4):
        if prune:
            eng.nl.set_prune(prune, float(sys_.cutoff) + 0.5 * float(sys_.skin))
        eng.frc.fill_(1.0e3)
        forces = []
        for _ in range(3):
            eng.force_step(0, 0, from_crd=True, overwrite=True)
            torch.cuda.synchronize()
            forces.append(eng.frc.cpu().numpy().copy())
        f = forces[-1]
        assert np.isfinite(f).all()
        assert np.array_equal(forces[1], f), "repeated steps at the same coordinates differ"
        if prune:
            # the pruning step sums its pairs in the order of the lists it
            # filters, the later steps in the inner lists': equal but for the
            # last bits
            np.testing.assert_allclose(forces[0], f, rtol=0, atol=1e-6 * float(np.abs(f).max()))
            assert np.array_equal(forces[0] == 0.0, f == 0.0)
        else:
            assert np.array_equal(forces[0], f), "repeated steps at the same coordinates differ"
        left = np.where(np.all(f == 0.0, axis=1))[0]
        say(f"    prune {prune}: atoms left out: {left.tolist()}")
        assert 1 <= len(left) <= 8
        keep = np.setdiff1d(np.arange(sys_.n), left)
        sub = _subsystem(sys_, keep)
        ref = oracle.full_reference(sub, beta=eng.pme.beta, K=(eng.pme.fftx, eng.pme.ffty, eng.pme.fftz))
        e = rel_rms(f[keep], ref["frc"])
        say(f"    slotted atoms vs the oracle without the left-out ones: force rel rms {e:.3e}")
        assert e < 1e-5


def _subsystem(sys_, keep):
    """The system restricted to atoms `keep` (exclusions renumbered)."""
    keep = np.asarray(keep)
    new = -np.ones(sys_.n, np.int64)
    new[keep] = np.arange(len(keep))
    rows = []
    for i in keep:
        s, c = int(sys_.excl_start[i]), int(sys_.excl_numbers[i])
        rows.append(sorted(int(new[j]) for j in sys_.excl_list[s:s + c] if new[j] >= 0))
    numbers = np.array([len(r) for r in rows], np.int32)
    start = np.zeros(len(keep), np.int32)
    start[1:] = np.cumsum(numbers)[:-1]
    lst = np.array([j for r in rows for j in r], np.int32)
    return dataclasses.replace(sys_, crd=sys_.crd[keep].copy(), charge=sys_.charge[keep].copy(),
                               mass=sys_.mass[keep].copy(), lj_type=sys_.lj_type[keep].copy(),
                               excl_start=start, excl_numbers=numbers, excl_list=lst, waters=0)


def test_grid_axis_above_256():
    """A 280-point PME axis (SPONGE's size rule for a ~275 A box length):
    the prep kernels' per-plane histograms hold it; forces and the
    reciprocal energy match the oracle. A 600-point axis is refused."""
    import torch
    from sponge_amd import nb
    sys_ = system.small_random(400, (275.0, 16.0, 15.0), seed=9, cutoff=5.0, skin=1.0)
    say(f"[case] long box {sys_.box.tolist()} n={sys_.n}")
    eng = nb.NonbondedEngine(sys_, refresh_interval=0)
    assert eng.pme.fftx == 280, eng.pme.fftx
    eng.update_neighbor_list(forced=1)
    eng.force_step(1, 1, from_crd=True, overwrite=True)
    torch.cuda.synchronize()
    ref = oracle.full_reference(sys_, beta=eng.pme.beta, K=(eng.pme.fftx, eng.pme.ffty, eng.pme.fftz))
    np.testing.assert_array_equal(eng.nl.pairs(), ref["pairs"])
    e = rel_rms(eng.frc.cpu().numpy(), ref["frc"])
    p = eng.energy_parts()
    say(f"    K {eng.pme.fftx}x{eng.pme.ffty}x{eng.pme.fftz}: force rel rms {e:.3e}, "
        f"reciprocal {p['reciprocal']:.9g} vs {ref['e_rec']:.9g}")
    assert e < 1e-5
    assert abs(p["reciprocal"] - ref["e_rec"]) <= 1e-6 * abs(ref["e_rec"])
    with pytest.raises(nb.SpongeError):
        nb.Particle_Mesh_Ewald(sys_.n, eng.box, sys_.cutoff, 1e-5, 600, 16, 16)


TOL_ABS = 5e-4   # kcal/mol/A: per-atom FP32 force error bound of this box


def test_prune_displacement_vote():
    """The inner list is pruned to rc + 0.5 A with a 20-step schedule. A pair
    8.5-8.6 A apart at the prune is not in it; one of its atoms then moves
    0.6 A towards the other (past half the prune margin, 0.25 A), bringing the
    pair inside rc = 8 A. The next fused step must re-prune on its own (the
    prep kernel's vote) and include the pair: forces equal the oracle's at the
    new coordinates, while leaving the pair out would move that atom's force
    by more than the tolerance. A move below the margin leaves the list."""
    import torch
    from sponge_amd import nb
    sys_ = system.config(0)
    say(f"[case] prune vote, n={sys_.n}")
    eng = nb.NonbondedEngine(sys_, refresh_interval=20)
    rp = sys_.cutoff + 0.5
    eng.nl.set_prune(20, rp)
    eng.update_neighbor_list(forced=1)
    eng.force_step(0, 0, from_crd=True, overwrite=True)   # prunes at these coordinates
    torch.cuda.synchronize()
    crd = eng.crd.cpu().numpy().astype(np.float64)
    L = sys_.box.astype(np.float64)
    oxy = np.arange(0, 3 * sys_.waters, 3)
    d = crd[oxy][:, None, :] - crd[oxy][None, :, :]
    d -= L * np.round(d / L)
    r = np.linalg.norm(d, axis=2)
    ii, jj = np.where((r > rp + 0.01) & (r < rp + 0.1))
    assert len(ii), "no oxygen pair just outside the prune radius"
    i, j = int(oxy[ii[0]]), int(oxy[jj[0]])
    v = crd[j] - crd[i]
    v -= L * np.round(v / L)
    u = v / np.linalg.norm(v)
    votes0 = eng.nl.prune_votes()
    # a move below the margin: no re-prune, forces still the oracle's
    moved = crd.copy()
    moved[i] = np.mod(crd[i] + 0.2 * u, L)
    eng.crd.copy_(torch.from_numpy(moved.astype(np.float32)))
    eng.force_step(0, 0, from_crd=True, overwrite=True)
    torch.cuda.synchronize()
    assert eng.nl.prune_votes() == votes0
    ref = oracle.full_reference(dataclasses.replace(sys_, crd=moved.astype(np.float32)), beta=eng.pme.beta,
                                K=(eng.pme.fftx, eng.pme.ffty, eng.pme.fftz))
    assert np.abs(eng.frc.cpu().numpy() - ref["frc"]).max() < TOL_ABS
    # past the margin: the pair comes inside the cutoff
    moved[i] = np.mod(crd[i] + 0.6 * u, L)
    s2 = dataclasses.replace(sys_, crd=moved.astype(np.float32))
    eng.crd.copy_(torch.from_numpy(moved.astype(np.float32)))
    eng.force_step(0, 0, from_crd=True, overwrite=True)
    torch.cuda.synchronize()
    assert eng.nl.prune_votes() == votes0 + 1
    ref = oracle.full_reference(s2, beta=eng.pme.beta, K=(eng.pme.fftx, eng.pme.ffty, eng.pme.fftz))
    f = eng.frc.cpu().numpy()
    err = np.abs(f - ref["frc"]).max()
    # what leaving the pair out would change on atom i
    pairs = ref["pairs"]
    keep = ~(((pairs[:, 0] == min(i, j)) & (pairs[:, 1] == max(i, j))))
    assert keep.sum() == len(pairs) - 1
    full = oracle.direct(s2, ref["uint"], pairs, ref["beta"])
    without = oracle.direct(s2, ref["uint"], pairs[keep], ref["beta"])
    miss = float(np.abs(full["frc"][i] - without["frc"][i]).max())
    say(f"    pair ({i}, {j}) at {np.linalg.norm(v):.3f} -> {np.linalg.norm(v) - 0.6:.3f} A: max force error "
        f"{err:.2e}, the pair's force on atom {i} {miss:.2e}")
    assert miss > 3 * TOL_ABS
    assert err < TOL_ABS


def test_two_level_outer_lists_from_the_rebuild():
    """Two-level per-slot lists (prune radius below the list radius): on a
    box with skin 0.5 A the outer lists hold the full list's pairs (within
    rc + 0.5 A at the rebuild), and the inner lists (rc + 0.2 A) are pruned
    from them at the steps' own positions. An oxygen pair 8.06-8.09 A apart
    at the rebuild is pulled apart by 0.45 A before the first fused step
    (8.51-8.54 A there: outer lists pruned at that step's positions would
    leave it out), then pushed 0.45 A towards the other atom from where it
    was at the rebuild (7.61-7.64 A: inside rc = 8 A). The second step
    re-prunes the inner lists on the displacement vote and must find the
    pair: forces equal the oracle's at both coordinates, and leaving the pair
    out would move the force on one of its atoms by more than the
    tolerance."""
    import torch
    from sponge_amd import nb
    sys_ = dataclasses.replace(system.config(0), skin=0.5)
    say(f"[case] two-level lists, n={sys_.n}, skin {sys_.skin}")
    eng = nb.NonbondedEngine(sys_, refresh_interval=20)
    eng.nl.set_prune(20, sys_.cutoff + 0.2)
    eng.update_neighbor_list(forced=1)
    torch.cuda.synchronize()
    crd = eng.crd.cpu().numpy().astype(np.float64)
    L = sys_.box.astype(np.float64)
    oxy = np.arange(0, 3 * sys_.waters, 3)
    d = crd[oxy][:, None, :] - crd[oxy][None, :, :]
    d -= L * np.round(d / L)
    r = np.linalg.norm(d, axis=2)
    ii, jj = np.where((r > 8.06) & (r < 8.09))
    assert len(ii), "no oxygen pair 8.06-8.09 A apart"
    i, j = int(oxy[ii[0]]), int(oxy[jj[0]])
    v = crd[j] - crd[i]
    v -= L * np.round(v / L)
    u = v / np.linalg.norm(v)
    votes0 = eng.nl.prune_votes()
    K = (eng.pme.fftx, eng.pme.ffty, eng.pme.fftz)
    errs = []
    for step, shift in enumerate((-0.45, 0.45)):
        moved = crd.copy()
        moved[i] = np.mod(crd[i] + shift * u, L)
        s2 = dataclasses.replace(sys_, crd=moved.astype(np.float32))
        eng.crd.copy_(torch.from_numpy(moved.astype(np.float32)))
        eng.force_step(0, 0, from_crd=True, overwrite=True)
        torch.cuda.synchronize()
        ref = oracle.full_reference(s2, beta=eng.pme.beta, K=K)
        f = eng.frc.cpu().numpy()
        errs.append(float(np.abs(f - ref["frc"]).max()))
        # the first step prunes on schedule; the second on the vote
        assert eng.nl.prune_votes() == votes0 + step
    pairs = ref["pairs"]
    keep = ~(((pairs[:, 0] == min(i, j)) & (pairs[:, 1] == max(i, j))))
    assert keep.sum() == len(pairs) - 1
    full = oracle.direct(s2, ref["uint"], pairs, ref["beta"])
    without = oracle.direct(s2, ref["uint"], pairs[keep], ref["beta"])
    miss = float(np.abs(full["frc"][[i, j]] - without["frc"][[i, j]]).max())
    say(f"    pair ({i}, {j}) at {np.linalg.norm(v):.3f} A: max force error {errs[0]:.2e} (first step), "
        f"{errs[1]:.2e} (voted re-prune); the pair's force {miss:.2e}")
    assert miss > 3 * TOL_ABS
    assert max(errs) < TOL_ABS


@pytest.mark.parametrize("schedule", ["tuned", "two-level"])
def test_pruned_list_under_motion(schedule):
    """The 23,570-atom box relaxed to ~300 K, then MD steps with the tuned
    default prune (sponge_nb_default_shapes, plus displacement votes), or
    with inner lists pruned every 4 steps to rc + 0.25 A from outer lists at
    rc + skin: at the last step before a refresh the fused step's forces
    equal a prune-off evaluation, after a forced rebuild, at the same
    coordinates."""
    import torch
    from sponge_amd import md, nb
    sys_ = system.config(1)
    say(f"[case] pruned list under motion, n={sys_.n}, {schedule}")
    mde = md.MDEngine(sys_, refresh_interval=20, temperature=300.0, seed=5)
    if schedule == "two-level":
        mde.nb.nl.set_prune(4, float(sys_.cutoff) + 0.25)
    mde.prepare()
    mde.equilibrate(600)
    for _ in range(20):
        mde.step()
    while mde.nb.steps_to_refresh() != 1:
        mde.step()
    mde.force()   # the 20th force step on this list
    torch.cuda.synchronize()
    f = mde.nb.frc.cpu().numpy()
    T = mde.temperature()
    votes = mde.nb.nl.prune_votes()
    crd = mde.nb.crd.cpu().numpy()
    e2 = nb.NonbondedEngine(dataclasses.replace(sys_, crd=crd), refresh_interval=20)
    e2.nl.set_prune(0)
    e2.update_neighbor_list(forced=1)
    e2.force_step(0, 0, from_crd=True, overwrite=True)
    torch.cuda.synchronize()
    g = e2.frc.cpu().numpy()
    e = rel_rms(f, g)
    say(f"    T {T:.1f} K, voted prunes {votes}, force rel rms vs prune-off {e:.2e}, "
        f"max abs {np.abs(f - g).max():.2e}")
    assert 200.0 < T < 450.0
    assert e < 1e-5
