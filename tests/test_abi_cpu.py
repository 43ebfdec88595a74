"""CPU checks of the C ABI library and the host-side helpers (no GPU calls)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from sponge_amd import system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, h)).read(), flags=re.S)
            names |= set(re.findall(r"\b(sponge_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from sponge_amd import nb
    path = nb.LIB_PATH
    if not os.path.exists(path):
        pytest.skip("libsponge_nb.so not built")
    lib = ctypes.CDLL(path)
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # the Python mirror binds exactly the header's functions
    assert sorted(nb.exported_symbols()) == names


def test_library_is_gfx950_code_object():
    from sponge_amd import nb
    if not os.path.exists(nb.LIB_PATH):
        pytest.skip("libsponge_nb.so not built")
    blob = open(nb.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_product_has_no_oracle_dependency():
    """The shipped package never imports or links the oracle."""
    pkg = os.path.join(ROOT, "sponge_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                src = open(os.path.join(dirpath, f), errors="replace").read()
                assert "import oracle" not in src and "from oracle" not in src, f
                assert "liboracle" not in src, f


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_water_box_generator():
    s = system.config(1)
    assert s.n == 23570 and s.waters == 7850
    assert (s.crd >= 0).all() and (s.crd < s.box).all()
    assert abs(float(s.charge.sum())) < 1e-2
    # TIP3P geometry
    o, h1, h2 = s.crd[0].astype(float), s.crd[1].astype(float), s.crd[2].astype(float)
    d = lambda a, b: np.linalg.norm(((a - b + s.box / 2) % s.box) - s.box / 2)
    assert d(o, h1) == pytest.approx(0.9572, abs=1e-5)
    assert d(o, h2) == pytest.approx(0.9572, abs=1e-5)
    # density of water ~1.0 g/cm^3
    mass = float(s.mass.astype(float).sum()) / 6.02214076e23
    vol = float(np.prod(s.box.astype(float))) * 1e-24
    assert mass / vol == pytest.approx(1.0, abs=0.02)
    # deterministic for a seed
    s2 = system.config(1)
    assert np.array_equal(s.crd, s2.crd)
    # exclusions: O->H1,H2 and H1->H2, upper triangular and sorted
    assert s.excl_numbers[:3].tolist() == [2, 1, 0]
    assert s.excl_list[:3].tolist() == [1, 2, 2]


def test_lj_tables_lorentz_berthelot():
    s = system.config(0)
    A, B = s.pair_table_square()
    sig, eps = 3.15061, 0.1521
    assert A[0, 0] / 12 == pytest.approx(4 * eps * sig ** 12, rel=1e-6)
    assert B[0, 0] / 6 == pytest.approx(4 * eps * sig ** 6, rel=1e-6)
    assert A[1, 1] == 0 and B[0, 1] == 0
    assert np.array_equal(A, A.T)


def test_pme_rejects_unsupported_fft_size():
    """A grid whose radix passes exceed the in-LDS line kernels is refused at
    creation (before any device call) instead of being transformed wrongly."""
    from sponge_amd import nb
    if not os.path.exists(nb.LIB_PATH):
        pytest.skip("libsponge_nb.so not built")
    lib = nb.load_library()
    h = ctypes.c_void_p()
    box = (ctypes.c_float * 3)(400.0, 400.0, 400.0)
    rc = lib.sponge_pme_create(1000, box, 10.0, 1e-5, 384, 384, 384, ctypes.byref(h))
    assert rc == -1
    assert b"384" in lib.sponge_nb_last_error()


def test_default_shapes_follow_the_tuned_table():
    """sponge_nb_default_shapes (a table lookup, no device call) returns the
    row of sponge_amd/csrc/tune_table.hpp that covers the atom count."""
    from sponge_amd import autotune, nb
    if not os.path.exists(nb.LIB_PATH):
        pytest.skip("libsponge_nb.so not built")
    lib = nb.load_library()
    rows = autotune.parse_table()
    assert rows and rows[-1]["max_atoms"] == 0x7fffffff
    assert [r["max_atoms"] for r in rows] == sorted(r["max_atoms"] for r in rows)
    for atoms in (100, 648, 5000, 23570, 50001, 96120, 10 ** 7):
        for refresh in (0, 10, 20):
            sh = nb.Shapes()
            assert lib.sponge_nb_default_shapes(atoms, refresh, 10.0, 2.0, ctypes.byref(sh)) == 0
            r = autotune.table_row(rows, atoms)
            assert sh.pair_waves == r["pair_waves"] and sh.conv_threads == r["conv_threads"]
            assert sh.conv_split == r["conv_split"] and sh.step_form == r["step_form"]
            want = max(1, refresh // r["prune_div"]) if refresh and r["prune_div"] else 0
            assert sh.prune_interval == want
            assert sh.prune_radius == pytest.approx(10.0 + r["prune_margin"] * 2.0)
    for r in rows:
        assert r["pair_waves"] in (2, 4, 8) and r["conv_threads"] in (256, 512)
        assert r["conv_split"] in (-1, 0, 1, 2) and r["step_form"] in (0, 1)
    sh = nb.Shapes()
    assert lib.sponge_nb_default_shapes(0, 20, 10.0, 2.0, ctypes.byref(sh)) != 0


def test_shape_candidates_cover_the_swept_axes():
    from sponge_amd import autotune
    c = autotune.candidates(20, 10.0, 2.0)
    assert len(c) == len(autotune.PAIR_WAVES) * len(autotune.CONV_FORMS) * 2 * len(autotune.PRUNES)
    assert {x["prune_interval"] for x in c} == {20, 10, 5, 4, 2, 0}
    assert all(x["prune_radius"] >= 10.0 for x in c)
    assert not any(x["conv_split"] == 0 for x in autotune.candidates(20, 10.0, 2.0, plane_fits=False))
