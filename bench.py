#!/usr/bin/env python3
"""ns/day of SPONGE's nonbonded force path on one MI355X.

Workload (default): the 23,570-atom PME water box (7,850 TIP3P waters +
10 Na+ + 10 Cl-, 61.9 A cubic box, rc 10 A, skin 2 A, PME 64^3 order 4,
2 fs) -- one step = main.cu's NVE step: uint coordinates, fused LJ + PME-direct
pair force, excluded-pair correction, PME reciprocal (spread, R2C FFT, BC
convolution, C2R FFT, gather), leap-frog with SETTLE rigid waters, and the
neighbor-list refresh every 20 steps, all on the GPU through libsponge_nb.so.
The trajectory starts from the seeded lattice with rigid-body velocities at
300 K, relaxed by untimed MD steps with velocities rescaled to 300 K.
Inputs are resident in HBM before timing.

Prints one JSON line (see the keys below). `--gpus N` with N > 1 exits:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 (vector) peak
PEAK_HBM_GBS = 8000.0
PAIR_FLOPS = 50              # per in-cutoff pair, force-only variant
# HBM bytes per launch of the force-only pair kernel, by workload: rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE of the same launches (KiB per dispatch,
# averaged over the run's pair launches), each scaled by what the counter
# reads for a 256 MiB device copy (FETCH x 2.0, WRITE x 1.0). Profile
# constants: a kernel change needs a new counter run.
PAIR_TRAFFIC_BYTES = {"water23570": 41.75e6, "water96120": 174.47e6}
# Algorithmic HBM bytes: pair kernel 44 per atom (16-byte record read, force
# 12 written, 16 of exclusion partners and list bookkeeping) + 4 per listed
# atom pair; spread 24 per atom (base, fraction + charge, bin entry) + 8 per
# half-spectrum point written; gather 44 per atom (base, fraction + charge,
# force read and written) + 4 per real grid point read.


def cpu_baseline(sys_, crd, vel, dt, refresh, beta, K, steps=200, warmup=20):
    """The OpenMP FP32 CPU engine of the same step (oracle/cpu_md.cpp: the
    reference's cell list, pair and PME formulas, its own FFT, leap-frog +
    SETTLE) on the host cores, from the GPU run's start state: `warmup`
    untimed steps, then `steps` timed."""
    import dataclasses
    import oracle

    s2 = dataclasses.replace(sys_, crd=crd)
    md = oracle.CpuMd(s2, vel, dt, refresh=refresh, beta=beta, K=K)
    md.run(warmup)
    t0 = time.perf_counter()
    md.run(steps)
    t_step = (time.perf_counter() - t0) / steps
    return {
        "value": round(0.1728 / t_step, 4),
        "unit": "ns/day",
        "cores": oracle.CpuMd.threads(),
        "kind": "port",
        "sample": (f"{steps} NVE steps after {warmup} warmup of the {sys_.n}-atom box (FP32 OpenMP engine: cell "
                   f"list every {refresh} steps, LJ + PME direct, exclusion correction, PME "
                   f"{K[0]}x{K[1]}x{K[2]} with its own FFT, leap-frog + SETTLE); {t_step * 1e3:.2f} ms/step"),
    }


def time_md(mde, steps, warmup, refresh, pme_events=False, graphs=True):
    """Times `steps` MD steps three ways: eager launches, one hipGraph per
    kind of step (plain / with the list rebuild), and one hipGraph holding a
    whole refresh period (refresh - 1 plain steps and the rebuilding one;
    steps beyond whole periods from the per-step graphs). The kernels are
    timed in a separate eager pass with every kernel launched on its own
    (serial form: the fused step's pair blocks ride in other launches), by
    events the launches record themselves. Returns per-step milliseconds."""
    import torch
    eng = mde.nb
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        mde.step()
    K = steps
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for k in range(K):
        mde.step()
    t_end.record(stream)
    torch.cuda.synchronize()
    out = {"eager": t_start.elapsed_time(t_end) / K}

    # kernel times, serial form
    mk = lambda: [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    ev = {"pair": (mk(), mk())}
    if pme_events:
        ev["spread"] = (mk(), mk())
        ev["gather"] = (mk(), mk())
    for a, b in ev.values():   # torch creates the hipEvent on first record
        for e in a + b:
            e.record(stream)
    mde.serial = True
    torch.cuda.synchronize()
    for k in range(K):
        eng.lj.set_pair_events(ev["pair"][0][k], ev["pair"][1][k])
        if pme_events:
            eng.pme.set_events((ev["spread"][0][k], ev["spread"][1][k]), (ev["gather"][0][k], ev["gather"][1][k]))
        mde.step()
    torch.cuda.synchronize()
    mde.serial = False
    eng.lj.set_pair_events(None, None)
    if pme_events:
        eng.pme.set_events(None, None)
    for name, (a, b) in ev.items():
        out[name + "_kernel"] = sum(x.elapsed_time(y) for x, y in zip(a, b)) / K

    def replay(fn, n):
        torch.cuda.synchronize()
        t_start.record(stream)
        for k in range(n):
            fn(k)
        t_end.record(stream)
        torch.cuda.synchronize()
        return t_start.elapsed_time(t_end)

    def graphs():
        g_plain, g_rebuild = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        with torch.cuda.graph(g_plain):
            mde.step(forced_update=0, stream=torch.cuda.current_stream())
        with torch.cuda.graph(g_rebuild):
            mde.step(forced_update=1, stream=torch.cuda.current_stream())
        g_block = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g_block):
            for j in range(refresh):
                mde.step(forced_update=1 if (j + 1) % refresh == 0 else 0, stream=torch.cuda.current_stream())
        return g_plain, g_rebuild, g_block

    def time_graphs(tag):
        g_plain, g_rebuild, g_block = graphs()
        per_step = lambda k: (g_rebuild if (k + 1) % refresh == 0 else g_plain).replay()
        for k in range(max(warmup, refresh)):
            per_step(k)
        out["graph" + tag] = replay(per_step, K) / K
        # the refresh-period graph, warmed up by two replays of its own
        for _ in range(2):
            g_block.replay()
        periods, rest = divmod(K, refresh)
        out["refresh_graph" + tag] = replay(lambda k: g_block.replay() if k < periods else per_step(k - periods),
                                            periods + rest) / K

    if not graphs:                       # eager only (profiling runs)
        out["graph"] = out["refresh_graph"] = out["eager"]
        out["md_steps_run"] = warmup + 2 * K
        return out
    time_graphs("")
    n_graph_steps = 2 * K + max(warmup, refresh) + 3 * refresh
    out["md_steps_run"] = warmup + 2 * K + n_graph_steps
    return out


def scale_bench(steps: int, warmup: int, refresh: int, equilibrate: int):
    """The 96,120-atom box (32,000 waters + 60 Na+ + 60 Cl-, 98.6 A, PME
    108^3): per-step time and the achieved bandwidth of its pair, spread and
    gather kernels (algorithmic bytes over launch-recorded kernel times)."""
    import torch
    from sponge_amd import md, system
    sys_ = system.config(2)
    mde = md.MDEngine(sys_, refresh_interval=refresh, temperature=300.0, seed=5)
    mde.prepare()
    mde.equilibrate(equilibrate)
    eng = mde.nb
    mde.force(need_energy=1, need_virial=1)
    pairs_in = eng.energy_parts()["pairs_in"]
    listed = eng.nl.status()["pairs"]
    shapes = eng.shapes().as_dict()
    t = time_md(mde, steps, warmup, refresh, pme_events=True)
    ms = min(t["eager"], t["graph"], t["refresh_graph"])
    n = sys_.n
    Kx, Ky, Kz = eng.pme.fftx, eng.pme.ffty, eng.pme.fftz
    b_pair = 44.0 * n + 4.0 * 2 * listed
    b_spread = 24.0 * n + 8.0 * Kx * Ky * (Kz // 2 + 1)
    b_gather = 44.0 * n + 4.0 * Kx * Ky * Kz
    gbs = lambda b, ms_: round(b / (ms_ * 1e-3) / 1e9, 1)
    temperature = mde.temperature()
    del mde, eng
    torch.cuda.empty_cache()
    return {
        "workload": f"{sys_.name}: {n} atoms, box {float(sys_.box[0])} A, rc 10 skin 2, PME {Kx}x{Ky}x{Kz}, "
                    f"NVE 2 fs with SETTLE, list every {refresh} steps",
        "steps": steps,
        "us_per_step": round(ms * 1e3, 2),
        "ns_per_day": round(0.1728 / (ms * 1e-3), 1),
        "eager_us_per_step": round(t["eager"] * 1e3, 2),
        "graph_us_per_step": round(t["graph"] * 1e3, 2),
        "refresh_graph_us_per_step": round(t["refresh_graph"] * 1e3, 2),
        "pairs_in_cutoff": int(pairs_in),
        "shapes": shapes,
        "pair_kernel_us": round(t["pair_kernel"] * 1e3, 2),
        "pair_gbs": gbs(b_pair, t["pair_kernel"]),
        "pair_tflops": round(PAIR_FLOPS * pairs_in / (t["pair_kernel"] * 1e-3) / 1e12, 2),
        "spread_kernel_us": round(t["spread_kernel"] * 1e3, 2),
        "spread_gbs": gbs(b_spread, t["spread_kernel"]),
        "gather_kernel_us": round(t["gather_kernel"] * 1e3, 2),
        "gather_gbs": gbs(b_gather, t["gather_kernel"]),
        "final_temperature_K": round(temperature, 1),
    }


def fep_batch_bench(steps: int, warmup: int, refresh: int):
    """Batched FEP: 16 soft-core lambda windows of one solvated 12-atom solute
    plus 8 replicas of 1.4k-12k atoms (cutoff 9 A, skin 2 A, PME ~1 A grids),
    every system taking whole NVE MD steps (force step, leap-frog + SETTLE,
    list update every `refresh` steps): (a) one system after another, each
    with its own tuned single-system step, issued eagerly and replayed from
    each system's own step graphs, (b) as one batch, one launch per stage,
    eagerly and (c) replayed from one hipGraph per MD step (one graph for
    steps with the list refresh, one without). Returns window-steps/s of each
    and the batch's force deviation from the single-system step at the same
    coordinates."""
    import torch
    from sponge_amd import md, system

    systems = system.fep_batch()
    engines = [md.MDEngine(s, refresh_interval=refresh, seed=7 + k) for k, s in enumerate(systems)]
    for e in engines:
        e.prepare()
        e.equilibrate(100)          # the lattice start's strain heat rescaled away
    # same accuracy: the batch's forces against each system's own step
    for e in engines:
        e.force()
    torch.cuda.synchronize()
    own = [e.nb.frc.cpu().numpy().astype(np.float64) for e in engines]
    bmd = md.BatchedMD(engines)
    bmd.prepare()
    bmd.batch.force_step(0, 0)
    torch.cuda.synchronize()
    dev = max(float(np.sqrt(((e.nb.frc.cpu().numpy() - f) ** 2).sum() / (f ** 2).sum()))
              for e, f in zip(engines, own))
    stream = torch.cuda.current_stream()

    def timed(fn, n):
        for k in range(warmup):
            fn(k)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for k in range(n):
            fn(k)
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) * 1e-3

    def seq_step(k):
        for e in engines:
            e.step(forced_update=1 if k % refresh == 0 else 0)

    for e in engines:
        e.prepare()
    t_seq = timed(seq_step, steps)
    # the same per-system steps replayed from each system's own two graphs,
    # one system after another (no host issue cost)
    seq_graphs = []
    for e in engines:
        gp, gr = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        with torch.cuda.graph(gp):
            e.step(stream=torch.cuda.current_stream(), forced_update=0)
        with torch.cuda.graph(gr):
            e.step(stream=torch.cuda.current_stream(), forced_update=1)
        seq_graphs.append((gp, gr))

    def seq_graph_step(k):
        for gp, gr in seq_graphs:
            (gr if k % refresh == 0 else gp).replay()

    t_seq_graph = timed(seq_graph_step, steps)
    bmd.prepare()

    def batch_step(k):
        bmd.step(forced_update=1 if k % refresh == 0 else 0)

    t_batch = timed(batch_step, steps)
    g_plain, g_rebuild = bmd.capture()

    def graph_step(k):
        (g_rebuild if k % refresh == 0 else g_plain).replay()

    t_graph = timed(graph_step, steps)
    temps = [e.temperature() for e in engines]
    nsys = len(systems)
    ws = lambda t: round(nsys * steps / t, 1)
    return {
        "metric": "window-steps/s (whole NVE MD steps), 16 soft-core FEP windows + 8 replicas",
        "systems": nsys,
        "atoms": int(sum(s.n for s in systems)),
        "sizes": [int(s.n) for s in systems],
        "steps": steps,
        "refresh": refresh,
        "sequential": ws(t_seq),
        "sequential_graph": ws(t_seq_graph),
        "batched": ws(t_batch),
        "batched_graph": ws(t_graph),
        "speedup_batched": round(t_seq / t_batch, 3),
        "speedup_batched_graph": round(t_seq / t_graph, 3),
        "speedup_batched_graph_vs_sequential_graph": round(t_seq_graph / t_graph, 3),
        "batch_vs_single_force_rel_rms_max": float(f"{dev:.3g}"),
        "final_temperature_K_range": [round(min(temps), 1), round(max(temps), 1)],
    }


def fep_frames_bench(frames: int, nwin: int = 16):
    """FEP / TI frame evaluation (main_fep.cu / main_ti.cu per frame, nonbonded
    part) of one solvated 12-atom solute in ~23.5k atoms at `nwin` lambda
    windows with charges and LJ perturbed, all inputs resident on the GPU:
    (a) the reference's way, one SPONGE_FEP evaluation per window (a soft-core
    handle and a charge set per window: Get_Energy_With_Coulomb_Direct plus
    the PME correction, self and reciprocal energies of that charge set from
    one pipeline), (b) every window in one
    pass (sponge_lj_soft_lambda_windows + sponge_pme_lambda_windows), and
    (c) one TI frame (dH/dlambda and the PME cross terms). Every frame
    rebuilds the neighbor list, as main_fep.cu:170 does. Returns
    frame-windows/s of (a) and (b) and frames/s of (c)."""
    import numpy as np
    import torch
    from sponge_amd import nb, system

    fs = system.fep_system(7850, 0.5, seed=11)
    eng = nb.NonbondedEngine(fs, refresh_interval=0)
    lams = np.linspace(0.0, 1.0, nwin)
    lam0 = float(np.float32(fs.lam))
    qA, qB = fs.charge_A.astype(np.float64), fs.charge_B.astype(np.float64)
    dq = torch.as_tensor((qB - qA).astype(np.float32), device="cuda")
    q_w = [torch.as_tensor(((1 - l) * qA + l * qB).astype(np.float32), device="cuda") for l in lams]
    soft = [nb.LJ_SOFT_CORE(fs.n, fs.ntypes, fs.ntypes_B, fs.lj_A, fs.lj_B, fs.lj_A_B, fs.lj_B_B, fs.lj_type,
                            fs.lj_type_B, fs.mask, float(l), fs.cutoff, eng.box, fs.sc_alpha, fs.sc_p,
                            fs.sc_sigma, fs.sc_sigma_min) for l in lams]
    P = eng.pme
    ex = (eng.excl_start, eng.excl_list, eng.excl_numbers)

    def rebuild():
        eng.update_neighbor_list(forced=1)
        eng.crd_to_uint()

    def per_window():
        rebuild()
        for w in range(nwin):
            soft[w].Get_Energy_With_Coulomb_Direct(eng.uint_crd, eng.nl, q_w[w], P.beta)
            P.lambda_windows(eng.uint_crd, q_w[w], None, eng.scaler, *ex, lams[w], [lams[w]])

    def all_windows():
        rebuild()
        eng.lj.lambda_windows(eng.nl, eng.uint_crd, eng.charge, dq, P.beta, lams)
        P.lambda_windows(eng.uint_crd, eng.charge, dq, eng.scaler, *ex, lam0, lams)

    def ti_frame():
        rebuild()
        eng.lj.Get_Partial_H_Partial_Lambda_With_Columb_Direct(eng.uint_crd, eng.charge, eng.nl, dq, 1, P.beta)
        P.cross_dhdl(eng.nl, eng.uint_crd, eng.charge, dq, eng.scaler, *ex, lj_pertubated=1)

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / frames

    t_ref = timed(per_window)
    t_all = timed(all_windows)
    t_ti = timed(ti_frame)
    for s in soft:
        s.Clear()
    return {
        "metric": "frame-windows/s, FEP partition energies (nonbonded) of a 23.5k-atom frame",
        "workload": f"{fs.name}: {fs.n} atoms, {nwin} lambda windows, charges and LJ perturbed, list rebuilt "
                    f"every frame, PME {P.fftx}x{P.ffty}x{P.fftz}",
        "frames": frames,
        "per_window_ms_per_frame": round(t_ref * 1e3, 3),
        "all_windows_ms_per_frame": round(t_all * 1e3, 3),
        "per_window_frame_windows_per_s": round(nwin / t_ref, 1),
        "all_windows_frame_windows_per_s": round(nwin / t_all, 1),
        "speedup_all_windows": round(t_ref / t_all, 2),
        "ti_ms_per_frame": round(t_ti * 1e3, 3),
    }


def adjoint_bench(eng, reps: int):
    """Config-5 timing: one forward force step, one energy-gradient call and
    one force-matching-gradient call (target forces from charges perturbed by
    N(0, 0.02) e, computed on the GPU), each averaged over `reps` launches."""
    import numpy as np
    import torch
    from sponge_amd import system

    stream = torch.cuda.current_stream()
    eng.update_neighbor_list(forced=1)
    P = len(eng.sys.lj_A)
    gq = torch.zeros(eng.n, dtype=torch.float64, device="cuda")
    gA = torch.zeros(P, dtype=torch.float64, device="cuda")
    gB = torch.zeros(P, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    q0 = eng.charge.clone()
    rng = np.random.default_rng(5)
    dq = torch.as_tensor(rng.normal(0.0, 0.02 * system.CHARGE_SCALE, eng.n), dtype=torch.float32, device="cuda")
    eng.charge.copy_(q0 + dq)
    eng.force_step(0, 0, from_crd=True, overwrite=True)
    target = eng.frc.clone()
    eng.charge.copy_(q0)
    eng.force_step(0, 0, from_crd=True, overwrite=True)
    frc = eng.frc.clone()

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(reps):
            fn()
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / reps

    fwd = timed(lambda: eng.force_step(0, 0, from_crd=True, overwrite=True))
    e_grad = timed(lambda: eng.energy_gradient(gq, gA, gB))
    l_grad = timed(lambda: eng.force_matching_gradient(frc, target, loss, gq, gA, gB))
    return {
        "workload": f"{eng.sys.name}: gradients w.r.t. {eng.n} charges and {P}+{P} LJ table entries",
        "forward_step_ms": round(fwd, 4),
        "energy_gradient_ms": round(e_grad, 4),
        "force_matching_gradient_ms": round(l_grad, 4),
        "energy_gradient_per_forward": round(e_grad / fwd, 3),
        "force_matching_gradient_per_forward": round(l_grad / fwd, 3),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", type=int, default=1, help="0: 648 atoms, 1: 23,570, 2: 96,120")
    ap.add_argument("--refresh", type=int, default=20)
    ap.add_argument("--equilibrate", type=int, default=2000,
                    help="untimed MD steps with velocities rescaled to 300 K before timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="run the reciprocal pipeline after the pair kernel on one stream")
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--adjoint-reps", type=int, default=50,
                    help="launches per timing of the gradient calls (0 skips them)")
    ap.add_argument("--fep-steps", type=int, default=200,
                    help="steps of the batched-FEP comparison (0 skips it)")
    ap.add_argument("--fep-frames", type=int, default=20,
                    help="frames of the FEP / TI frame-evaluator comparison (0 skips it)")
    ap.add_argument("--scale-steps", type=int, default=100,
                    help="steps of the 96,120-atom measurement (0 skips it)")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep the launch shapes on this workload first and time the fastest (default: the "
                         "shapes the tuned table gives this atom count)")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager launches only (for counter collection, which serialises graph replays)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (calibrates HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {args.gpus} is not supported (it would shard "
              "by spatial decomposition: bricks of cell columns with a cutoff + skin halo)", file=sys.stderr)
        sys.exit(2)

    import torch
    from sponge_amd import md, system

    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4, "launches": 5}))
        return

    sys_ = system.config(args.config)
    mde = md.MDEngine(sys_, refresh_interval=args.refresh, temperature=300.0, seed=5)
    mde.serial = args.serial
    eng = mde.nb

    # Main_Initial: forced rebuild; then the synthetic lattice start is
    # relaxed (velocities rescaled to 300 K every step) so that the timed
    # trajectory is liquid water near 300 K
    mde.prepare()
    mde.equilibrate(args.equilibrate)
    shapes_source = "tuned table (sponge_amd/csrc/tune_table.hpp) for this atom count"
    if args.autotune:
        from sponge_amd import autotune
        us, _, res = autotune.tune(mde, args.refresh, log=lambda *x: print(*x, file=sys.stderr, flush=True))
        shapes_source = f"online sweep of {len(res)} candidates ({us:.2f} us/step)"
    t_start_K = mde.temperature()
    crd0 = eng.crd.cpu().numpy().copy()
    vel0 = mde.vel.cpu().numpy().copy()
    mde.force(need_energy=1, need_virial=1)
    pairs_in0 = eng.energy_parts()["pairs_in"]
    nl_pairs = eng.nl.status()["pairs"]

    t = time_md(mde, args.steps, args.warmup, args.refresh, graphs=not args.no_graph)
    eager_ms, graph_ms, block_ms, pair_ms = t["eager"], t["graph"], t["refresh_graph"], t["pair_kernel"]
    # the fastest of the ways of issuing the same steps
    ms = min(eager_ms, graph_ms, block_ms)

    eng.nl.check()
    eng.update_neighbor_list(forced=1)
    mde.force(need_energy=1, need_virial=1)
    pairs_in1 = eng.energy_parts()["pairs_in"]
    pairs_in = 0.5 * (pairs_in0 + pairs_in1)
    temperature = mde.temperature()

    flops = PAIR_FLOPS * pairs_in
    achieved = flops / (pair_ms * 1e-3) / 1e12
    traffic = PAIR_TRAFFIC_BYTES.get(sys_.name)
    out = {
        "metric": "ns/day (24k-atom PME water, 2 fs, NVE MD step with SETTLE)",
        "value": round(0.1728 / (ms * 1e-3), 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 6),
        "eager_ms_per_step": round(eager_ms, 6),
        "graph_ms_per_step": round(graph_ms, 6),
        "refresh_graph_ms_per_step": round(block_ms, 6),
        "timed_as": ("eager" if ms == eager_ms else "graph per step" if ms == graph_ms else
                     f"graph of {args.refresh} steps") + (", every kernel on its own" if args.serial else
                                                          ", pair kernel on a second stream"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"{sys_.name}: {sys_.n} atoms, box {float(sys_.box[0])} A, rc {sys_.cutoff} "
                         f"skin {sys_.skin}, PME {eng.pme.fftx}x{eng.pme.ffty}x{eng.pme.fftz} order 4, "
                         f"NVE leap-frog 2 fs with SETTLE rigid waters, neighbor list every "
                         f"{args.refresh} steps"),
            "atoms": sys_.n,
            "listed_pairs": int(nl_pairs),
            "pairs_in_cutoff": int(pairs_in),
            "equilibration_steps": args.equilibrate,
            "start_temperature_K": round(t_start_K, 1),
            "md_steps_run": t["md_steps_run"],
            "final_temperature_K": round(temperature, 1),
            "parallelism": "single GPU",
            "shapes": eng.shapes().as_dict(),
            "shapes_source": shapes_source,
        },
        "roofline": {
            "kernel": "pair_kernel<false> (LJ + PME direct), timed launched on its own",
            "bound": "valu",
            "note": "FP32 VALU bound (no MFMA shape); peak = MI355X FP32 vector 157.3 TFLOP/s",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "traffic_unit": "bytes per launch",
            "traffic_source": ("profile constant: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload, "
                               "scaled by a 256 MiB copy" if traffic else None),
            "kernel_ms": round(pair_ms, 6),
            "flops_per_launch": flops,
        },
    }
    del mde, eng
    torch.cuda.empty_cache()
    if args.scale_steps > 0:
        out["scale"] = scale_bench(args.scale_steps, 20, args.refresh, max(500, args.equilibrate // 2))
    if args.adjoint_reps > 0:
        from sponge_amd import nb
        out["adjoint"] = adjoint_bench(nb.NonbondedEngine(sys_, refresh_interval=args.refresh), args.adjoint_reps)
    if args.fep_steps > 0:
        out["fep_batch"] = fep_batch_bench(args.fep_steps, 20, args.refresh)
    if args.fep_frames > 0:
        out["fep_frames"] = fep_frames_bench(args.fep_frames)
    if not args.no_cpu_baseline:
        import oracle
        beta = float(np.float32(oracle.pme_beta(sys_.cutoff, sys_.pme_tolerance)))
        K = tuple(sys_.fft[d] if sys_.fft[d] > 0 else system.fft_size(sys_.box[d]) for d in range(3))
        out["cpu_baseline"] = cpu_baseline(sys_, crd0, vel0, 0.002 * system.TIME_UNIT_PER_PS, args.refresh, beta,
                                           K, steps=args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
