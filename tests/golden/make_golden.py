"""Writes the committed golden vectors of the 648-atom water box (seed 2024).

    python tests/golden/make_golden.py

The FP64 oracle produces them: forces (total and per part), the energy parts,
the LJ virial, the per-atom energy arrays, and the neighbor list as per-atom
neighbor counts plus a SHA-256 of the sorted (i<j) pair array. They pin the
oracle against regressions and give the GPU tests fixed expected values.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from sponge_amd import system  # noqa: E402


def pair_digest(pairs):
    p = np.ascontiguousarray(pairs, dtype=np.int32)
    return hashlib.sha256(p.tobytes()).hexdigest()


def make(sys_):
    ref = oracle.full_reference(sys_)
    pairs = ref["pairs"]
    counts = np.bincount(pairs[:, 0], minlength=sys_.n) + np.bincount(pairs[:, 1], minlength=sys_.n)
    arrays = dict(
        crd=sys_.crd.astype(np.float32), uint=ref["uint"].astype(np.uint32),
        frc=ref["frc"], frc_direct=ref["frc_direct"], frc_excl=ref["frc_excl"],
        frc_rec=ref["frc_rec"], atom_energy=ref["atom_energy"], atom_virial=ref["atom_virial"],
        atom_direct=ref["atom_direct"], atom_corr=ref["atom_corr"],
        neighbor_counts=counts.astype(np.int32))
    scalars = dict(
        name=sys_.name, n=sys_.n, box=[float(b) for b in sys_.box], cutoff=sys_.cutoff,
        skin=sys_.skin, beta=float(ref["beta"]), K=list(ref["K"]),
        pair_count=int(len(pairs)), pair_sha256=pair_digest(pairs), n_in=int(ref["n_in"]),
        e_lj=ref["e_lj"], e_direct=ref["e_direct"], w_lj=ref["w_lj"], e_corr=ref["e_corr"],
        e_rec=ref["e_rec"], e_self=ref["e_self"], e_lrc=ref["e_lrc"])
    return arrays, scalars


if __name__ == "__main__":
    arrays, scalars = make(system.config(0))
    np.savez_compressed(os.path.join(HERE, "water648.npz"), **arrays)
    with open(os.path.join(HERE, "water648.json"), "w") as f:
        json.dump(scalars, f, indent=1)
    print({k: v for k, v in scalars.items() if k != "pair_sha256"})
