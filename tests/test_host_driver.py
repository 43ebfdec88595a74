"""The C++ host mirror and the step-loop driver (sponge_amd/sponge_nb_driver).

CPU tests run the driver's `check_input` mode, which parses the mdin and the
input files and touches no device: they pin the mdin rules of control.cu
(blocks, comments, commas, `main` / `comments` blocks, duplicate keys,
default_in_file_prefix, command-line flags) and the file formats. The GPU test
runs the SPONGE-ordered step loop on SPONGE-format files and checks the forces
and the printed energies against the FP64 oracle, and its NVE mode (leap-frog
with SETTLE) against the oracle's trajectory, with the trajectory and restart
files it writes.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from sponge_amd import system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sponge_amd", "sponge_nb_driver")

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def run(args, cwd=None, timeout=120):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)


def commands(stdout):
    return dict(re.findall(r"^command (\S+) = (.*)$", stdout, flags=re.M))


@needs_driver
def test_mdin_rules(tmp_path):
    sys_ = system.config(0)
    system.write_sponge_inputs(sys_, str(tmp_path / "in"))
    (tmp_path / "pre_charge.txt").write_text((tmp_path / "in" / "charge.txt").read_text())
    mdin = tmp_path / "mdin.txt"
    mdin.write_text(
        "block test\n"
        "cutoff = 8.0, skin = 1.5\n"
        "  neighbor_list\n  {\n    refresh_interval = 5   # trailing text\n    skin_permit = 0.6\n  }\n"
        "# comment = 9, hidden = 3\n"
        "##disabled { cutoff = 1.0\n  skin = 9 }\n"
        "PME { fftx = 24, ffty=30 }\n"
        "main { step_limit = 7 }\n"
        "comments { anything = 1 }\n"
        f"coordinate_in_file = {tmp_path / 'in' / 'coordinate.txt'}\n"
        f"default_in_file_prefix = {tmp_path / 'pre'}\n"
        "check_input = 1\n")
    r = run(["-mdin", str(mdin), "-extra", "3", "4"])
    assert r.returncode == 0, r.stderr
    c = commands(r.stdout)
    assert r.stdout.splitlines()[0] == "block test"
    assert c["cutoff"] == "8.0" and c["skin"] == "1.5"
    assert c["neighbor_list_refresh_interval"] == "5"
    assert c["neighbor_list_skin_permit"] == "0.6"
    assert c["PME_fftx"] == "24" and c["PME_ffty"] == "30"
    assert c["step_limit"] == "7"                      # `main` block adds no prefix
    assert c["extra"] == "3"                           # value = first token of `3 4`
    assert c["charge_in_file"] == str(tmp_path / "pre") + "_charge.txt"
    for hidden in ("comment", "hidden", "disabled_cutoff", "anything", "comments_anything"):
        assert hidden not in c
    assert "atoms 648 " in r.stdout and "skin 1.5 steps 7" in r.stdout


@needs_driver
def test_duplicate_key_is_an_error(tmp_path):
    mdin = tmp_path / "mdin.txt"
    mdin.write_text("dup\ncutoff = 8.0\nskin = 1\ncutoff = 9.0\n")
    r = run(["-mdin", str(mdin), "-check_input", "1"])
    assert r.returncode != 0
    assert "cutoff is set more than once" in r.stderr


@needs_driver
def test_input_files_round_trip(tmp_path):
    sys_ = system.small_random(300, (12.0, 13.0, 14.0), seed=4, cutoff=4.5, skin=1.0,
                               fft=(12, 14, 15))
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), extra={"check_input": 1})
    r = run(["-mdin", mdin])
    assert r.returncode == 0, r.stderr
    m = re.search(r"atoms (\d+) box (\S+) (\S+) (\S+) excluded (\d+) charge_sum (\S+)", r.stdout)
    assert m, r.stdout
    assert int(m.group(1)) == sys_.n
    np.testing.assert_allclose([float(m.group(i)) for i in (2, 3, 4)], sys_.box, rtol=1e-6)
    assert int(m.group(5)) == len(sys_.excl_list)
    assert abs(float(m.group(6)) - float(sys_.charge.astype(np.float64).sum())) < 1e-3
    c = commands(r.stdout)
    assert c["PME_fftx"] == "12" and c["PME_fftz"] == "15"


@needs_driver
def test_only_nve_and_npt_are_accepted(tmp_path):
    sys_ = system.config(0)
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), velocities=np.zeros((sys_.n, 3)),
                                      extra={"mode": "NVT"})
    r = run(["-mdin", mdin])
    assert r.returncode != 0 and "NVE, or NPT" in r.stderr
    # NPT needs one of the two barostats, and no thermostat
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), velocities=np.zeros((sys_.n, 3)),
                                      extra={"mode": "NPT"})
    r = run(["-mdin", mdin])
    assert r.returncode != 0 and "barostat" in r.stderr
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), velocities=np.zeros((sys_.n, 3)),
                                      extra={"mode": "NPT", "barostat": "berendsen_barostat",
                                             "thermostat": "middle_langevin"})
    r = run(["-mdin", mdin])
    assert r.returncode != 0 and "thermostat" in r.stderr


@needs_driver
def test_md_inputs_parse(tmp_path):
    sys_ = system.config(0)
    v = system.rigid_velocities(sys_)
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), velocities=v, dt_ps=0.002,
                                      extra={"check_input": 1})
    r = run(["-mdin", mdin])
    assert r.returncode == 0, r.stderr
    c = commands(r.stdout)
    assert c["mode"] == "NVE" and c["dt"] == "0.002"
    lines = (tmp_path / "constrain.txt").read_text().split("\n")
    assert int(lines[0]) == 3 * sys_.waters and lines[1].split()[:2] == ["0", "1"]


@needs_driver
def test_malformed_exclusions_are_rejected(tmp_path):
    sys_ = system.config(0)
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), extra={"check_input": 1})
    ex = tmp_path / "exclude.txt"
    lines = ex.read_text().splitlines()
    lines[1] = "2 1 99999"                             # atom index out of range
    ex.write_text("\n".join(lines) + "\n")
    r = run(["-mdin", mdin])
    assert r.returncode != 0 and "malformed exclude_in_file" in r.stderr


def _rst7_case(tmp_path, irest=None, with_vel=True, extra=None):
    """A 300-atom system written as an rst7 with a few atoms placed outside
    the box, read by check_input and exported again as an rst7."""
    sys_ = system.small_random(300, (12.0, 13.0, 14.0), seed=4, cutoff=4.5, skin=1.0, fft=(12, 14, 15))
    L = sys_.box.astype(np.float64)
    crd = sys_.crd.astype(np.float64).copy()
    crd[0, 0] = -1.5                       # one box length below: +L
    crd[1, 1] = L[1] + 0.75                # past the box: -L
    crd[2, 2] = -2.0 * L[2] + 0.5          # two box lengths below: two passes
    crd[3, 0] = L[0]                       # exactly L maps to 0
    vel = np.random.default_rng(7).normal(0.0, 0.01, (sys_.n, 3)) if with_vel else None
    sys_.crd = crd.astype(np.float32)
    keys = {"check_input": 1, "rst": str(tmp_path / "out.rst7")}
    if irest is not None:
        keys["amber_irest"] = irest
    keys.update(extra or {})
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), extra=keys, velocities=vel, rst7=True)
    if vel is not None:                    # write_sponge_inputs writes the time as 0.0: give it one
        system.write_rst7(str(tmp_path / "system.rst7"), sys_.name, sys_.crd, sys_.box, vel, 12.5)
    r = run(["-mdin", mdin])
    return sys_, vel, r


@needs_driver
def test_rst7_read_and_export_round_trip(tmp_path):
    """amber_rst7 (MD_core.cu:1258-1347) read with its velocities and time,
    every coordinate wrapped into [0, L) by the periodic map, then written
    back in the rst7 layout of Export_Restart_File (MD_core.cu:1425-1463):
    title, `%8d %.3lf`, two atoms a line of %12.7f, box and angles."""
    sys_, vel, r = _rst7_case(tmp_path)
    assert r.returncode == 0, r.stderr
    assert "residues 300 start_time 12.500 amber_irest 1" in r.stdout
    title, n, time, crd, v, box = system.read_rst7(str(tmp_path / "out.rst7"))
    assert title == sys_.name and n == sys_.n and time == 12.5
    L = sys_.box.astype(np.float64)
    _, _, _, crd_in, _, _ = system.read_rst7(str(tmp_path / "system.rst7"))
    want = crd_in.copy()
    for _ in range(10):                     # Crd_Periodic_Map, ten passes
        want = np.where(want < 0, want + L, np.where(want < L, want, want - L))
    assert (crd >= 0).all() and (crd < L).all()
    np.testing.assert_allclose(crd, want, atol=2e-6)
    assert abs(crd[0, 0] - (L[0] - 1.5)) < 1e-5 and abs(crd[1, 1] - 0.75) < 1e-5
    assert abs(crd[2, 2] - 0.5) < 1e-5 and crd[3, 0] == 0.0
    np.testing.assert_allclose(v, vel, atol=6e-8)
    np.testing.assert_allclose(box, L, atol=1e-6)
    lines = (tmp_path / "out.rst7").read_text().splitlines()
    assert lines[1] == f"{sys_.n:8d} 12.500"
    assert len(lines) == 2 + 2 * (sys_.n // 2) + 1
    assert len(lines[2]) == 72 and lines[-1].endswith("  90.0000000  90.0000000  90.0000000")


@needs_driver
@pytest.mark.parametrize("irest,with_vel", [(0, True), (1, False)])
def test_rst7_zero_velocities(tmp_path, irest, with_vel):
    """amber_irest 0, or an rst7 with no time (so no velocities): every
    velocity is zero and, without velocities, the start time too."""
    sys_, _, r = _rst7_case(tmp_path, irest=irest, with_vel=with_vel)
    assert r.returncode == 0, r.stderr
    assert "All velocity will be set to 0" in r.stdout
    _, n, time, _, v, _ = system.read_rst7(str(tmp_path / "out.rst7"))
    assert (v == 0).all() and v.shape == (sys_.n, 3)
    assert time == (12.5 if with_vel else 0.0)


@needs_driver
def test_rst7_negative_irest_exports_text(tmp_path):
    """amber_irest -1 keeps the velocities but writes the text restart pair
    (MD_core.cu:1464-1487)."""
    sys_, vel, r = _rst7_case(tmp_path, irest=-1)
    assert r.returncode == 0, r.stderr
    rst = str(tmp_path / "out.rst7")
    head = open(rst + "_coordinate.txt").readline().split()
    assert head == [str(sys_.n), "12.500"]
    np.testing.assert_allclose(np.loadtxt(rst + "_velocity.txt", skiprows=1), vel, atol=6e-8)


@needs_driver
def test_rst7_truncated_is_an_error(tmp_path):
    sys_, _, r = _rst7_case(tmp_path)
    lines = (tmp_path / "system.rst7").read_text().splitlines()
    (tmp_path / "system.rst7").write_text("\n".join(lines[:100]) + "\n")
    r = run(["-mdin", str(tmp_path / "mdin.txt")])
    assert r.returncode != 0 and "amber_rst7 ends early" in r.stderr


@needs_driver
def test_residue_file(tmp_path):
    """residue_in_file (MD_core.cu:784-826): `natom nres`, then each
    residue's atom count; counts past natom are refused."""
    sys_ = system.config(0)
    (tmp_path / "res.txt").write_text(f"{sys_.n} {sys_.waters}\n" + "3\n" * sys_.waters)
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), extra={"check_input": 1,
                                                                  "residue_in_file": str(tmp_path / "res.txt")})
    r = run(["-mdin", mdin])
    assert r.returncode == 0, r.stderr
    assert f"residues {sys_.waters} " in r.stdout
    (tmp_path / "res.txt").write_text(f"{sys_.n} 2\n600\n100\n")
    r = run(["-mdin", mdin])
    assert r.returncode != 0 and "malformed residue_in_file" in r.stderr


DRIVER_CASES = {
    "water648": lambda: system.config(0),
    "random_noncubic": lambda: system.small_random(900, (21.0, 24.5, 27.3), seed=3, cutoff=7.0,
                                                   skin=1.5),
}


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("name", sorted(DRIVER_CASES))
def test_driver_step_loop_matches_oracle(tmp_path, name):
    import oracle
    sys_ = DRIVER_CASES[name]()
    print(f"[driver] {name} n={sys_.n}", flush=True)
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), step_limit=3, write_mdout_interval=1,
                                      extra={"frc_out_file": str(tmp_path / "frc.txt"),
                                             "mdout": str(tmp_path / "mdout.txt")})
    r = run(["-mdin", mdin], timeout=100)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    beta = float(re.search(r"PME beta ([0-9.eE+-]+)", r.stdout).group(1))
    K = tuple(int(v) for v in re.search(r"fft (\d+) (\d+) (\d+)", r.stdout).groups())
    ref = oracle.full_reference(sys_, beta=np.float32(beta), K=K)
    frc = np.loadtxt(tmp_path / "frc.txt", skiprows=1)
    err = float(np.sqrt(((frc - ref["frc"]) ** 2).sum() / (ref["frc"] ** 2).sum()))
    assert err < 1e-5, err
    steps = re.findall(r"^step (\d+) LJ (\S+) PME (\S+) reciprocal (\S+) self (\S+) direct (\S+) "
                       r"correction (\S+)$", r.stdout, flags=re.M)
    assert [int(s[0]) for s in steps] == [1, 2, 3]
    for s in steps:                                    # coordinates do not move
        e_lj, e_pme, e_rec, e_self, e_dir, e_corr = (float(v) for v in s[1:])
        pme_ref = ref["e_rec"] + ref["e_self"] + ref["e_direct"] + ref["e_corr"]
        scale = abs(ref["e_rec"]) + abs(ref["e_self"]) + abs(ref["e_direct"]) + abs(ref["e_corr"])
        assert abs(e_pme - pme_ref) < 1e-5 * scale
        assert abs(e_rec - ref["e_rec"]) < 1e-5 * scale
        assert abs(e_lj - (ref["e_lj"] + ref["e_lrc"])) < 1e-5 * abs(ref["e_lj"]) + 1e-5
    assert "overflow 0" in r.stdout
    # mdout (control.cu:538-573): the header, then one row per print step;
    # potential = atom energies (LJ + direct) + LRC + PME correction, self
    # and reciprocal (MD_core.cu:1800-1812), i.e. LJ + PME
    head, rows = read_mdout(tmp_path / "mdout.txt")
    assert head == ["step", "time", "temperature", "potential", "LJ", "PME"]
    assert [int(r_["step"]) for r_ in rows] == [1, 2, 3]
    total = ref["e_lj"] + ref["e_lrc"] + ref["e_rec"] + ref["e_self"] + ref["e_direct"] + ref["e_corr"]
    for r_, st in zip(rows, steps):
        assert r_["time"] == f"{0.001 * int(r_['step']):.3f}" and r_["temperature"] == "0.00"
        assert r_["LJ"] == f"{float(st[1]):.2f}" and r_["PME"] == f"{float(st[2]):.2f}"
        assert abs(float(r_["potential"]) - total) < 0.006 + 1e-5 * abs(total), (r_["potential"], total)
    assert re.search(r"^ +step = +1,  +time = +0.001,  temperature = +0.00, $", r.stdout, flags=re.M)


def read_mdout(path):
    lines = open(path).read().splitlines()
    head = lines[0].split()
    return head, [dict(zip(head, line.split())) for line in lines[1:]]


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("fmt", ["txt", "rst7"])
def test_driver_nve_matches_oracle(tmp_path, fmt):
    """SPONGE's NVE loop on the 648-atom rigid-water box, 10 steps of 2 fs:
    the restart coordinates after step 10 against the oracle's FP64
    trajectory from the same velocities (1e-4 Angstrom RMS, the bound of
    the chaotic growth of FP32 rounding over 10 steps), every water rigid,
    one trajectory frame per write_information_interval. The inputs are the
    text files or one AMBER rst7 (then the restart is an rst7 too, and the
    waters are the residues of the mdout temperature)."""
    import oracle
    sys_ = system.config(0)
    v = system.rigid_velocities(sys_)
    print(f"[driver] nve water648 10 steps from {fmt}", flush=True)
    extra = {"write_information_interval": 5, "crd": str(tmp_path / "mdcrd.dat"),
             "box": str(tmp_path / "mdbox.txt"), "rst": str(tmp_path / "restart"),
             "mdout": str(tmp_path / "mdout.txt")}
    if fmt == "rst7":
        (tmp_path / "res.txt").write_text(f"{sys_.n} {sys_.waters}\n" + "3\n" * sys_.waters)
        extra["residue_in_file"] = str(tmp_path / "res.txt")
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), step_limit=10, write_mdout_interval=5,
                                      refresh_interval=20, velocities=v, dt_ps=0.002, extra=extra,
                                      rst7=fmt == "rst7")
    if fmt == "rst7":                                  # the oracle starts from what the file holds
        _, _, _, c0, v0, _ = system.read_rst7(str(tmp_path / "system.rst7"))
        sys_.crd = c0.astype(np.float32)
        v = v0.astype(np.float32)
    r = run(["-mdin", mdin], timeout=100)
    sys.stdout.write(r.stdout[-1500:])
    assert r.returncode == 0, r.stderr
    assert "rigid triangle numbers is 216" in r.stdout
    frames = np.fromfile(tmp_path / "mdcrd.dat", dtype=np.float32).reshape(-1, sys_.n, 3)
    assert frames.shape[0] == 2
    assert len((tmp_path / "mdbox.txt").read_text().splitlines()) == 2
    if fmt == "txt":
        head = (tmp_path / "restart_coordinate.txt").read_text().splitlines()
        assert head[0].split()[0] == str(sys_.n) and abs(float(head[0].split()[1]) - 0.02) < 1e-9
        crd = np.loadtxt(tmp_path / "restart_coordinate.txt", skiprows=1, max_rows=sys_.n)
        vel = np.loadtxt(tmp_path / "restart_velocity.txt", skiprows=1)
        nres, res_of = sys_.n, np.arange(sys_.n)
    else:
        title, n, time, crd, vel, box = system.read_rst7(str(tmp_path / "restart"))
        assert title == sys_.name and n == sys_.n and abs(time - 0.02) < 1e-9
        np.testing.assert_allclose(box, sys_.box, atol=1e-6)
        nres, res_of = sys_.waters, np.arange(sys_.n) // 3
    np.testing.assert_allclose(crd, frames[-1], atol=1e-6)
    # the temperature column: residue centre-of-mass kinetic energy
    # (MD_core.cu:216-239, 1738-1742) of the step-10 velocities
    m = sys_.mass.astype(np.float64)
    p_res = np.zeros((nres, 3))
    np.add.at(p_res, res_of, m[:, None] * vel)
    m_res = np.bincount(res_of, weights=m, minlength=nres)
    t_ref = float((0.5 * (p_res ** 2).sum(1) / m_res).sum() * 2.0 / 0.00198716 / nres / 3.0)
    head_, rows = read_mdout(tmp_path / "mdout.txt")
    assert head_ == ["step", "time", "temperature", "potential", "LJ", "PME"]
    assert [r_["step"] for r_ in rows] == ["5", "10"] and rows[1]["time"] == "0.020"
    assert abs(float(rows[1]["temperature"]) - t_ref) < 0.006 + 1e-4 * t_ref, (rows[1]["temperature"], t_ref)
    groups = oracle.settle_groups(sys_.n, *system.water_constraints(sys_), sys_.mass)
    c_ref, v_ref, _ = oracle.md_run(sys_, v, 10, 0.002 * system.TIME_UNIT_PER_PS, groups)
    L = sys_.box.astype(np.float64)
    d = crd - c_ref
    d -= L * np.round(d / L)
    rms = float(np.sqrt((d ** 2).mean()))
    print(f"    RMS |dx| {rms:.3e}, vel rel RMS "
          f"{float(np.sqrt(((vel - v_ref) ** 2).sum() / (v_ref ** 2).sum())):.3e}", flush=True)
    assert rms < 1e-4
    x = crd.reshape(-1, 3, 3)
    oh = np.linalg.norm((x[:, 1:] - x[:, :1]) - L * np.round((x[:, 1:] - x[:, :1]) / L), axis=2)
    assert np.abs(oh - system.TIP3P_ROH).max() < 1e-4


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("barostat,keys", [
    ("berendsen_barostat", {"berendsen_barostat_tau": 0.05, "berendsen_barostat_update_interval": 5}),
    ("andersen_barostat", {"andersen_barostat_tau": 0.2})])
def test_driver_npt_matches_oracle(tmp_path, barostat, keys):
    """mode NPT with the Berendsen or the Andersen barostat over the NVE
    leap-frog, 10 steps of 2 fs on the 648-atom box: the pressure of the
    coupled steps, Main_Volume_Change and its list check; the box and the
    restart coordinates against the oracle's loop (box 1e-6 relative,
    positions 1e-4 Angstrom RMS)."""
    import oracle
    sys_ = system.config(0)
    v = system.rigid_velocities(sys_)
    print(f"[driver] npt {barostat} water648 10 steps", flush=True)
    extra = {"mode": "NPT", "barostat": barostat, "target_pressure": 1.0, "write_information_interval": 10,
             "crd": str(tmp_path / "mdcrd.dat"), "box": str(tmp_path / "mdbox.txt"),
             "rst": str(tmp_path / "restart"), "mdout": str(tmp_path / "mdout.txt")}
    extra.update(keys)
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), step_limit=10, write_mdout_interval=5,
                                      refresh_interval=0, velocities=v, dt_ps=0.002, extra=extra)
    r = run(["-mdin", mdin], timeout=100)
    sys.stdout.write(r.stdout[-800:])
    assert r.returncode == 0, r.stderr
    box = np.array([float(x) for x in (tmp_path / "mdbox.txt").read_text().split()[:3]])
    crd = np.loadtxt(tmp_path / "restart_coordinate.txt", skiprows=1, max_rows=sys_.n)
    groups = oracle.settle_groups(sys_.n, *system.water_constraints(sys_), sys_.mass)
    kind = "andersen" if barostat.startswith("andersen") else "berendsen"
    kw = dict(tau=keys[f"{barostat}_tau"])
    if kind == "berendsen":
        kw["update_interval"] = keys["berendsen_barostat_update_interval"]
    c_ref, _, b_ref, log = oracle.barostat_run(sys_, v, 10, 0.002, groups, kind=kind, **kw)
    d = crd - c_ref
    d -= b_ref * np.round(d / b_ref)
    rms = float(np.sqrt((d ** 2).mean()))
    print(f"    box {box[0]:.6f} / {b_ref[0]:.6f}, RMS |dx| {rms:.3e}", flush=True)
    assert abs(box[0] / b_ref[0] - 1.0) < 1e-6 + 1e-6 / b_ref[0]   # mdbox.txt prints 6 decimals
    assert abs(b_ref[0] / float(sys_.box[0]) - 1.0) > 1e-5
    assert rms < 1e-4
    # the barostat's mdout columns (Berendsen_barostat.cu:70-76,
    # andersen_barostat.cu:46-53): density of the step-10 box
    # (MD_core.cu:1510-1514) and the pressure the step printed
    head, rows = read_mdout(tmp_path / "mdout.txt")
    assert head == ["step", "time", "temperature", "potential", "LJ", "PME", "density", "pressure"] + \
        (["dV/dt"] if kind == "andersen" else [])
    rho = float(sys_.mass.astype(np.float64).sum()) * 1e24 / 6.023e23 / float(np.prod(b_ref))
    assert abs(float(rows[-1]["density"]) - rho) < 6e-5, (rows[-1]["density"], rho)
    p10 = float(re.search(r"^step 10 pressure (\S+) bar", r.stdout, flags=re.M).group(1))
    assert abs(float(rows[-1]["pressure"]) - p10) < 0.006
