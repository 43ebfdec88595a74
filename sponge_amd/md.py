"""Python mirror of the step on either side of the nonbonded force path:
SPONGE's NVE leap-frog (MD_INFORMATION::NVE_iteration::Leap_Frog,
MD_core/MD_core.cu:1535-1547) and the SETTLE module (constrain/SETTLE.cuh),
over include/sponge_md.h. `MDEngine` runs the reference's main loop
(main.cu:36-41: Main_Calculate_Force, then Main_Iteration) for NVE with rigid
waters: the fused nonbonded force step, one fused leap-frog + SETTLE launch,
and the neighbor-list update, all on the GPU with no host synchronisation.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import system
from .nb import _FP3, _I, _F, _P, NonbondedEngine, _check, _ptr, _stream, _vec3, load_library


def leap_frog(atom_numbers, vel, crd, frc, acc, inverse_mass, dt, max_velocity=0.0, stream=None):
    """MD_Iteration_Leap_Frog (MD_core/MD_core.cu:13-45)."""
    _check(load_library().sponge_md_leap_frog(int(atom_numbers), _ptr(vel), _ptr(crd), _ptr(frc), _ptr(acc),
                                              _ptr(inverse_mass), float(dt), float(max_velocity),
                                              _stream(stream)), "Leap_Frog")


class SETTLE:
    """constrain/SETTLE.cuh:37-67 over sponge_settle_*; built from the
    constraint pairs CONSTRAIN holds (atom_i, atom_j, constant_r)."""

    def __init__(self, atom_numbers, pair_i, pair_j, pair_r, h_mass, dt, exp_gamma=1.0, is_minimization=0):
        lib = load_library()
        pi = np.ascontiguousarray(pair_i, np.int32)
        pj = np.ascontiguousarray(pair_j, np.int32)
        pr = np.ascontiguousarray(pair_r, np.float32)
        m = np.ascontiguousarray(h_mass, np.float32)
        h = _P()
        _check(lib.sponge_settle_create(int(atom_numbers), len(pi), pi.ctypes.data_as(_P), pj.ctypes.data_as(_P),
                                        pr.ctypes.data_as(_P), m.ctypes.data_as(_P), float(dt), float(exp_gamma),
                                        int(is_minimization), ctypes.byref(h)), "SETTLE::Initial")
        self.h = h
        self.atom_numbers = int(atom_numbers)
        t, p, r = _I(0), _I(0), _I(0)
        _check(lib.sponge_settle_counts(h, ctypes.byref(t), ctypes.byref(p), ctypes.byref(r)), "settle counts")
        self.triangle_numbers, self.pair_numbers, self.remaining = t.value, p.value, r.value

    def groups(self):
        """Triangles (A, B, C), their canonical geometry, pairs and lengths."""
        nt, npr = self.triangle_numbers, self.pair_numbers
        tri = np.zeros((max(nt, 1), 3), np.int32)
        geo = np.zeros((max(nt, 1), 5), np.float32)
        prs = np.zeros((max(npr, 1), 2), np.int32)
        plen = np.zeros(max(npr, 1), np.float32)
        _check(load_library().sponge_settle_groups(self.h, tri.ctypes.data_as(_P), geo.ctypes.data_as(_P),
                                                   prs.ctypes.data_as(_P), plen.ctypes.data_as(_P)),
               "settle groups")
        return dict(tri=tri[:nt], geo=geo[:nt], pairs=prs[:npr], plen=plen[:npr], rest=self.remaining)

    def set_factors(self, v_factor, x_factor):
        _check(load_library().sponge_settle_set_factors(self.h, float(v_factor), float(x_factor)), "settle factors")

    def Remember_Last_Coordinates(self, uint_crd, scaler, stream=None):
        _check(load_library().sponge_settle_remember_last_coordinates(self.h, _ptr(uint_crd), _vec3(scaler),
                                                                      _stream(stream)),
               "Remember_Last_Coordinates")

    def Do_SETTLE(self, d_mass, crd, box_length, vel, need_pressure=0, d_pressure=None, stream=None):
        _check(load_library().sponge_settle_do(self.h, _ptr(d_mass), _ptr(crd), _vec3(box_length), _ptr(vel),
                                               int(need_pressure), _ptr(d_pressure), _stream(stream)), "Do_SETTLE")

    def Clear(self):
        if self.h:
            load_library().sponge_settle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.Clear()
        except Exception:
            pass


def md_iteration(settle, atom_numbers, vel, crd, frc, acc, inverse_mass, mass, uint_crd, box_length, dt,
                 max_velocity=0.0, need_pressure=0, pressure=None, stream=None):
    """Remember + Leap_Frog + Do_SETTLE as one launch (sponge_md_iteration)."""
    _check(load_library().sponge_md_iteration(
        settle.h if settle is not None else None, int(atom_numbers), _ptr(vel), _ptr(crd), _ptr(frc), _ptr(acc),
        _ptr(inverse_mass), _ptr(mass), _ptr(uint_crd), _vec3(box_length), float(dt), float(max_velocity),
        int(need_pressure), _ptr(pressure), _stream(stream)), "sponge_md_iteration")


class MDEngine:
    """NVE molecular dynamics of a `system.System` with rigid waters.

    One step is the reference's Main_Calculate_Force + Main_Iteration: the
    fused nonbonded force from the current coordinates (uint conversion
    fused, forces assigned), the fused leap-frog + SETTLE, and
    Neighbor_List_Update. Velocities are in Angstrom per internal time unit
    (1/20.455 ps), as SPONGE stores them."""

    def __init__(self, sys_, vel=None, dt_ps=0.002, refresh_interval=20, temperature=300.0, seed=5,
                 max_velocity=0.0, device="cuda"):
        import torch
        self.nb = NonbondedEngine(sys_, refresh_interval=refresh_interval, device=device)
        self.sys = sys_
        self.n = sys_.n
        self.dt = float(dt_ps) * system.TIME_UNIT_PER_PS
        self.max_velocity = float(max_velocity)
        if vel is None:
            vel = system.rigid_velocities(sys_, temperature, seed)
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
        self.vel = t(np.asarray(vel, np.float32).reshape(-1, 3), torch.float32).contiguous()
        m = sys_.mass.astype(np.float32)
        self.mass = t(m, torch.float32)
        self.inverse_mass = t(np.where(m > 0, 1.0 / np.where(m > 0, m, 1), 0).astype(np.float32), torch.float32)
        self.acc = torch.zeros((self.n, 3), dtype=torch.float32, device=device)
        self.pressure = torch.zeros(1, dtype=torch.float32, device=device)
        self.settle = None
        if sys_.waters:
            pi, pj, pr = system.water_constraints(sys_)
            self.settle = SETTLE(self.n, pi, pj, pr, m, self.dt)
        self.steps = 0
        # True: the reciprocal pipeline runs on the caller's stream after the
        # pair kernel instead of concurrently on the PME handle's stream
        self.serial = False

    @property
    def crd(self):
        return self.nb.crd

    def prepare(self):
        """Main_Initial's forced list build."""
        self.nb.update_neighbor_list(forced=1)

    def force(self, need_energy=0, need_virial=0, stream=None):
        self.nb.force_step(need_energy, need_virial, from_crd=True, overwrite=True, serial=self.serial,
                           stream=stream)

    def iterate(self, need_pressure=0, stream=None):
        md_iteration(self.settle, self.n, self.vel, self.nb.crd, self.nb.frc, self.acc, self.inverse_mass,
                     self.mass, self.nb.uint_crd, self.nb.box, self.dt, self.max_velocity, need_pressure,
                     self.pressure, stream)

    def step(self, need_energy=0, stream=None, forced_update=None):
        """One MD step. forced_update None: the list's own interval / skin
        rule; 1: rebuild now; 0: never (used for graph capture)."""
        self.force(need_energy, 0, stream)
        self.iterate(stream=stream)
        if forced_update is None:
            self.nb.update_neighbor_list(stream=stream)
        elif forced_update:
            self.nb.update_neighbor_list(forced=1, stream=stream)
        self.steps += 1

    def kinetic_energy(self):
        """Sum of m v^2 / 2 of the (half-step) velocities, FP64 on the host."""
        import torch
        torch.cuda.synchronize()
        v = self.vel.double().cpu().numpy()
        return 0.5 * float((self.sys.mass.astype(np.float64)[:, None] * v * v).sum())

    def potential_energy(self):
        p = self.nb.energy_parts()
        return p["lj"] + p["lrc"] + p["direct"] + p["correction"] + p["self"] + p["reciprocal"]
