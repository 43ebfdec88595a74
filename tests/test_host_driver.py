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
                                      extra={"frc_out_file": str(tmp_path / "frc.txt")})
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


@pytest.mark.gpu
@needs_driver
def test_driver_nve_matches_oracle(tmp_path):
    """SPONGE's NVE loop on the 648-atom rigid-water box, 10 steps of 2 fs:
    the restart coordinates after step 10 against the oracle's FP64
    trajectory from the same velocity file (1e-4 Angstrom RMS, the bound of
    the chaotic growth of FP32 rounding over 10 steps), every water rigid,
    one trajectory frame per write_information_interval."""
    import oracle
    sys_ = system.config(0)
    v = system.rigid_velocities(sys_)
    print("[driver] nve water648 10 steps", flush=True)
    mdin = system.write_sponge_inputs(sys_, str(tmp_path), step_limit=10, write_mdout_interval=5,
                                      refresh_interval=20, velocities=v, dt_ps=0.002,
                                      extra={"write_information_interval": 5,
                                             "crd": str(tmp_path / "mdcrd.dat"),
                                             "box": str(tmp_path / "mdbox.txt"),
                                             "rst": str(tmp_path / "restart")})
    r = run(["-mdin", mdin], timeout=100)
    sys.stdout.write(r.stdout[-1500:])
    assert r.returncode == 0, r.stderr
    assert "rigid triangle numbers is 216" in r.stdout
    frames = np.fromfile(tmp_path / "mdcrd.dat", dtype=np.float32).reshape(-1, sys_.n, 3)
    assert frames.shape[0] == 2
    assert len((tmp_path / "mdbox.txt").read_text().splitlines()) == 2
    head = (tmp_path / "restart_coordinate.txt").read_text().splitlines()
    assert head[0].split()[0] == str(sys_.n) and abs(float(head[0].split()[1]) - 0.02) < 1e-9
    crd = np.loadtxt(tmp_path / "restart_coordinate.txt", skiprows=1, max_rows=sys_.n)
    vel = np.loadtxt(tmp_path / "restart_velocity.txt", skiprows=1)
    np.testing.assert_allclose(crd, frames[-1], atol=1e-6)
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
             "rst": str(tmp_path / "restart")}
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
