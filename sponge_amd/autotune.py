"""Shape sweep of the fused MD step: the autotuner of the launch shapes.

SPONGE fixes its launch shapes in the source (Lennard_Jones_force.cu:1327-1350
`thread_LJ`, neighbor_list.cu:435, 480). Here every shape the fused step
exposes (include/sponge_nb.h, sponge_shapes) is chosen by measurement:

* pair_waves: i-clusters (one wave each) per block of the force-only pair
  kernel, 2 / 4 / 8;
* the convolution: one 256- or 512-thread workgroup per kz plane, the three
  line launches, or the slab form whose spread and inverse transform take the
  y transforms of an x plane each (conv_split 0 / 1 / 2);
* step_form: the pair kernel on a second stream or every kernel in turn;
* the inner list: pruned every refresh / d steps to cutoff + m * skin.

`sweep` times whole NVE MD steps of a running system for every candidate: a
hipGraph holding one refresh period (the timed form of bench.py) is captured
per candidate and replayed, and the per-step time is the minimum over a few
rounds visited in alternating order (so drifts of the running system do not
favour one end of the list). `tune` applies the fastest. Run as a module it
sweeps the water boxes of the benchmark configurations, writes the measured
table (markdown) and regenerates the compiled-in defaults
(sponge_amd/csrc/tune_table.hpp) that the handles' creators read:

    python -m sponge_amd.autotune --configs 0,1,2 --md sweep.md --table sponge_amd/csrc/tune_table.hpp
"""
import argparse
import itertools
import os
import sys
import time

PAIR_WAVES = (2, 4, 8)
CONV_FORMS = ((0, 256), (0, 512), (1, 256), (2, 256))  # (conv_split, conv_threads)
STEP_FORMS = (0, 1)
# (prune_div, margin / skin): pruned every refresh / prune_div steps to
# cutoff + margin * skin. A margin of 1 keeps every listed pair, per slot,
# from the step after each rebuild (one level: no pruning steps but that
# one, whatever the divisor); below 1 the inner lists are pruned from those
# (two levels); (0, 0) reads the cluster-pair list every step.
PRUNES = ((1, 1.0), (5, 0.5), (4, 0.5), (4, 0.75), (2, 0.5), (5, 0.25), (10, 0.25), (0, 0.0))

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "csrc", "tune_table.hpp")


def candidates(refresh, cutoff, skin, plane_fits=True):
    """Every combination of the swept shapes, as sponge_shapes dicts (with the
    prune divisor and margin they came from)."""
    out = []
    for w, (split, nt), form, (div, margin) in itertools.product(PAIR_WAVES, CONV_FORMS, STEP_FORMS, PRUNES):
        if split == 0 and not plane_fits:
            continue
        interval = max(1, refresh // div) if (div > 0 and refresh > 0) else 0
        out.append({"pair_waves": w, "conv_split": split, "conv_threads": nt, "step_form": form,
                    "prune_interval": interval, "prune_radius": float(cutoff + margin * skin),
                    "prune_div": div, "prune_margin": margin})
    return out


def _shape_fields(c):
    return {k: c[k] for k in ("pair_waves", "conv_split", "conv_threads", "step_form", "prune_interval",
                              "prune_radius")}


def time_candidate(mde, refresh, shapes, periods=4):
    """us per MD step of one candidate: set the shapes, capture one refresh
    period as a graph, warm it up with one replay, time `periods` replays."""
    import torch
    eng = mde.nb
    eng.set_shapes(_shape_fields(shapes))
    # one eager step first: whatever a shape needs on first use is made
    # outside the capture
    mde.step()
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for j in range(refresh):
            mde.step(forced_update=1 if (j + 1) % refresh == 0 else 0, stream=torch.cuda.current_stream())
    g.replay()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(periods):
        g.replay()
    t1.record(stream)
    torch.cuda.synchronize()
    us = t0.elapsed_time(t1) * 1e3 / (periods * refresh)
    del g
    return us


def sweep(mde, refresh, rounds=2, periods=4, cands=None, log=None):
    """Times every candidate `rounds` times (alternating order); returns
    [(us_per_step, candidate)] sorted, fastest first. The system keeps moving
    through the sweep; its list is rebuilt at the end of every period."""
    from sponge_amd import nb
    eng = mde.nb
    if cands is None:
        cands = candidates(refresh, float(eng.sys.cutoff), float(eng.sys.skin), plane_fits=_plane_fits(eng))
    best = [float("inf")] * len(cands)
    for r in range(rounds):
        order = range(len(cands)) if r % 2 == 0 else reversed(range(len(cands)))
        for i in order:
            try:
                us = time_candidate(mde, refresh, cands[i], periods)
            except nb.SpongeError as e:         # a shape the grid cannot take
                if log:
                    log(f"    skip {cands[i]}: {e}")
                continue
            best[i] = min(best[i], us)
        if log:
            log(f"  round {r + 1}/{rounds}: best so far {min(best):.2f} us/step")
    res = sorted(zip(best, cands), key=lambda t: t[0])
    return [(us, c) for us, c in res if us < float("inf")]


def _plane_fits(eng):
    """Whether the one-workgroup-per-plane convolution fits this grid (probed
    by setting it; the handle's shapes are restored)."""
    from sponge_amd import nb
    cur = eng.shapes()
    probe = nb.Shapes.from_dict(cur.as_dict())
    probe.conv_split = 0
    try:
        eng.set_shapes(probe)
        ok = True
    except nb.SpongeError:
        ok = False
    eng.set_shapes(cur)
    return ok


def tune(mde, refresh, rounds=2, periods=4, log=None):
    """Sweeps and applies the fastest candidate; returns (us, candidate, results)."""
    res = sweep(mde, refresh, rounds, periods, log=log)
    us, best = res[0]
    mde.nb.set_shapes(_shape_fields(best))
    return us, best, res


# ------------------------------------------------------------------ the table
def parse_table(path=TABLE_PATH):
    """Rows of kTuneTable in tune_table.hpp as dicts."""
    import re
    txt = open(path).read()
    body = txt[txt.index("kTuneTable[] = {"):]
    body = body[:body.index("};")]
    rows = []
    for m in re.finditer(r"\{\s*(0x[0-9a-fA-F]+|\d+),\s*(\d+),\s*(-?\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*([0-9.]+)f\}",
                         body):
        rows.append({"max_atoms": int(m.group(1), 0), "pair_waves": int(m.group(2)),
                     "conv_split": int(m.group(3)), "conv_threads": int(m.group(4)),
                     "step_form": int(m.group(5)), "prune_div": int(m.group(6)),
                     "prune_margin": float(m.group(7))})
    return rows


def table_row(rows, atoms):
    for r in rows:
        if atoms <= r["max_atoms"]:
            return r
    return rows[-1]


def write_table(winners, path=TABLE_PATH):
    """winners: [(atoms, grid, candidate, us, default_plane_rule)] sorted by
    atoms; a row covers sizes up to the geometric mean of its system's and
    the next system's atom counts."""
    lines = []
    for k, (atoms, grid, c, us, rule_split) in enumerate(winners):
        if k + 1 < len(winners):
            bound = int((atoms * winners[k + 1][0]) ** 0.5)
            b = f"{bound:10d}"
        else:
            b = "0x7fffffff"
        split = -1 if c["conv_split"] == rule_split else c["conv_split"]
        lines.append(f"    {{{b}, {c['pair_waves']}, {split:2d}, {c['conv_threads']}, {c['step_form']}, "
                     f"{c['prune_div']}, {c['prune_margin']}f}},  // {atoms} atoms, PME {grid}: {us:.2f} us/step")
    txt = open(TABLE_PATH).read()
    head = txt[:txt.index("constexpr TuneRow kTuneTable[] = {")]
    tail = txt[txt.index("// clang-format on"):]
    new = head + "constexpr TuneRow kTuneTable[] = {\n    // generated by sponge_amd/autotune.py\n" + \
        "\n".join(lines) + "\n};\n" + tail
    with open(path, "w") as f:
        f.write(new)


def _fmt(c):
    conv = {1: "split", 2: "slab"}.get(c["conv_split"], f"plane/{c['conv_threads']}")
    prune = "off" if c["prune_div"] == 0 else f"/{c['prune_div']} +{c['prune_margin']}skin"
    return f"| {c['pair_waves']} | {conv} | {'serial' if c['step_form'] else 'two streams'} | {prune} |"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--configs", default="0,1,2", help="benchmark water boxes (system.config indices)")
    ap.add_argument("--refresh", type=int, default=20)
    ap.add_argument("--equilibrate", type=int, default=300)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--periods", type=int, default=4)
    ap.add_argument("--md", default=None, help="write the measured table here (markdown)")
    ap.add_argument("--table", default=None, help="regenerate this tune_table.hpp from the winners")
    ap.add_argument("--top", type=int, default=8)
    a = ap.parse_args(argv)
    import torch
    from sponge_amd import md, system
    log = lambda *x: print(*x, flush=True)
    winners, sections = [], []
    for ci in [int(x) for x in a.configs.split(",") if x]:
        s = system.config(ci)
        log(f"[autotune] {s.name}: {s.n} atoms")
        mde = md.MDEngine(s, refresh_interval=a.refresh, temperature=300.0, seed=5)
        mde.prepare()
        mde.equilibrate(a.equilibrate)
        eng = mde.nb
        grid = f"{eng.pme.fftx}x{eng.pme.ffty}x{eng.pme.fftz}"
        default = eng.shapes().as_dict()
        t0 = time.time()
        res = sweep(mde, a.refresh, a.rounds, a.periods, log=log)
        log(f"  {len(res)} candidates in {time.time() - t0:.1f} s; fastest {res[0][0]:.2f} us/step")
        # the plane-size rule's choice (one workgroup per plane up to 64 x 64),
        # written as -1 in the table
        rule_split = 1 if (eng.pme.fftx * eng.pme.ffty > 64 * 64 or not _plane_fits(eng)) else 0
        us, best = res[0]
        winners.append((s.n, grid, best, us, rule_split))
        dflt = next(((u, k) for k, (u, c) in enumerate(res) if _shape_fields(c) == default), None)
        rows = [f"| {k + 1} | {u:.2f} " + _fmt(c) for k, (u, c) in enumerate(res[:a.top])]
        sec = [f"## {s.name}: {s.n} atoms, PME {grid}, list every {a.refresh} steps, 300 K",
               "",
               f"{len(res)} candidates, {a.rounds} rounds of {a.periods} refresh-period graph replays each "
               f"(minimum per candidate). The defaults compiled in before this sweep ranked "
               + (f"{dflt[1] + 1} ({dflt[0]:.2f} us/step)." if dflt else "outside the swept set."),
               "",
               "| rank | us/step | pair waves | convolution | step form | prune |",
               "|---|---|---|---|---|---|", *rows,
               f"| last | {res[-1][0]:.2f} " + _fmt(res[-1][1]), ""]
        sections.append("\n".join(sec))
        del mde, eng
        torch.cuda.empty_cache()
    winners.sort(key=lambda t: t[0])
    if a.md:
        head = ["# Launch-shape sweep of the fused NVE MD step (water boxes)", "",
                "Command: `python -m sponge_amd.autotune " + " ".join(argv if argv is not None else sys.argv[1:]) +
                "`", "",
                "Each candidate: one refresh period captured as a hipGraph and replayed; us per MD step "
                "(force step + leap-frog/SETTLE + the list rebuild every period). Pair waves: i-clusters per "
                "block of the force-only pair kernel; convolution: one workgroup (256 / 512 threads) per kz "
                "plane or the three line launches; prune: inner list every refresh/d steps to "
                "cutoff + m skin (off: the full list every step).", ""]
        with open(a.md, "w") as f:
            f.write("\n".join(head + sections))
    if a.table:
        write_table(winners, a.table)
    for atoms, grid, c, us, _ in winners:
        log(f"[autotune] {atoms} atoms: {us:.2f} us/step with {_shape_fields(c)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
