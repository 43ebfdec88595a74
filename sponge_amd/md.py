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
from .nb import _FP3, _I, _F, _P, BatchedSystems, NonbondedEngine, _check, _ptr, _stream, _vec3, load_library


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


def get_pressure(atom_numbers, vel, mass, atom_virial, virial, volume, pressure, ek=None, stream=None):
    """MD_INFORMATION::system_information::Get_Pressure (MD_core/MD_core.cu:1777-1795)."""
    _check(load_library().sponge_md_get_pressure(int(atom_numbers), _ptr(vel), _ptr(mass), _ptr(atom_virial),
                                                 _ptr(virial), float(volume), _ptr(pressure), _ptr(ek),
                                                 _stream(stream)), "Get_Pressure")


def pressure_from_parts(atom_numbers, vel, mass, parts, volume, pressure, ek=None, stream=None):
    """Get_Pressure from the fused step's FP64 parts (sponge_md_pressure_from_parts)."""
    _check(load_library().sponge_md_pressure_from_parts(int(atom_numbers), _ptr(vel), _ptr(mass), _ptr(parts),
                                                        float(volume), _ptr(pressure), _ptr(ek), _stream(stream)),
           "pressure_from_parts")


def scale_list(t, factor, stream=None):
    """Scale_List (common.cu): t *= factor in place (float32 device tensor)."""
    _check(load_library().sponge_md_scale_list(int(t.numel()), _ptr(t), float(factor), _stream(stream)),
           "Scale_List")


PRES_CONVERTION = 6.946827162543585e4          # bar per kcal/mol/A^3 (common.cuh:24-25)
PRES_CONVERTION_INVERSE = 0.00001439506089041446


def _f32(x):
    return float(np.float32(x))


class BERENDSEN_BAROSTAT_INFORMATION:
    """barostats/Berendsen_barostat.cu:5-88 and its step (main.cu:334-354):
    every update_interval steps, with the pressure p of the step,
    mu = 1 - update_interval * compressibility * dt / tau / 3 * (p0 - p)
    (dt and tau in ps, compressibility in bar^-1 converted to
    (kcal/mol/A^3)^-1), coordinates scaled by mu, then the volume change.
    The stochastic term (stochastic_term = 1) adds
    sqrt(2 kB T compressibility / tau / V) / 3 * N(0, dt_internal) to mu and
    scales the velocities by 1 / mu."""

    def __init__(self, target_pressure_bar, box_length, dt_ps=1e-3, tau=1.0, compressibility=4.5e-5,
                 update_interval=None, stochastic_term=0, random_seed=0, target_temperature=300.0):
        self.target_pressure = _f32(target_pressure_bar * PRES_CONVERTION_INVERSE)
        b = np.asarray(box_length, np.float32)
        self.V0 = _f32(b[0] * b[1] * b[2])
        self.newV = self.V0
        self.dt = _f32(dt_ps)
        self.taup = _f32(tau)
        self.compressibility = _f32(_f32(compressibility) * np.float32(PRES_CONVERTION))
        self.stochastic_term = int(stochastic_term)
        self.update_interval = int(update_interval) if update_interval else (1 if self.stochastic_term else 10)
        if self.stochastic_term and self.update_interval != 1:
            raise ValueError("the update interval should be 1 when using stochastic term")
        self.target_temperature = float(target_temperature)
        self.rng = np.random.default_rng(random_seed)
        self.is_initialized = 1

    def Ask_For_Calculate_Pressure(self, steps):
        return self.is_initialized and steps % self.update_interval == 0

    def scale_factor(self, p_now):
        """crd_scale_factor of main.cu:337-344 (FP32 like the reference)."""
        f = np.float32
        mu = f(1) - f(self.update_interval) * f(self.compressibility) * f(self.dt) / f(self.taup) / f(3) * \
            (f(self.target_pressure) - f(p_now))
        vel_scale = None
        if self.stochastic_term:
            noise = f(self.rng.normal(0.0, self.update_interval * self.dt * 20.455))
            mu = mu + f(np.sqrt(f(2) * f(system.KB) * f(self.target_temperature) * f(self.compressibility) /
                                f(self.taup) / f(self.newV))) / f(3) * noise
            vel_scale = float(f(1) / mu)
        return float(mu), vel_scale


class ANDERSEN_BAROSTAT_INFORMATION:
    """barostats/andersen_barostat.cu:5-57 and its step (main.cu:356-372):
    every step dV/dt += (p - p0) / W with 1/W = V0 compressibility / tau^2
    (tau in internal time units), V' = V + dV/dt dt, mu = cbrt(V' / V);
    coordinates scaled by mu, velocities by 1 / mu, then the volume change."""

    def __init__(self, target_pressure_bar, box_length, tau=1.0, compressibility=4.5e-5, dV_dt=0.0):
        self.target_pressure = _f32(target_pressure_bar * PRES_CONVERTION_INVERSE)
        b = np.asarray(box_length, np.float32)
        self.V0 = _f32(b[0] * b[1] * b[2])
        self.new_V = self.V0
        taup = _f32(_f32(tau) * np.float32(system.TIME_UNIT_PER_PS))
        comp = _f32(_f32(compressibility) * np.float32(PRES_CONVERTION))
        self.h_mass_inverse = _f32(np.float32(self.V0) * np.float32(comp) / np.float32(taup) / np.float32(taup))
        self.dV_dt = _f32(dV_dt)
        self.update_interval = 1
        self.is_initialized = 1

    def Ask_For_Calculate_Pressure(self, steps):
        return bool(self.is_initialized)

    def scale_factor(self, p_now, volume, dt_internal):
        f = np.float32
        self.dV_dt = float(f(self.dV_dt) + (f(p_now) - f(self.target_pressure)) * f(self.h_mass_inverse))
        self.new_V = float(f(volume) + f(self.dV_dt) * f(dt_internal))
        mu = float(np.cbrt(np.float64(self.new_V) / np.float64(volume)))
        return mu, float(f(1.0) / f(mu))


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
        self.barostat = None     # BERENDSEN_ / ANDERSEN_BAROSTAT_INFORMATION
        self.h_pressure = 0.0
        self.ek = torch.zeros(1, dtype=torch.float32, device=device)

    @property
    def crd(self):
        return self.nb.crd

    def prepare(self):
        """Main_Initial's forced list build; raises if it overflowed a
        capacity."""
        self.nb.update_neighbor_list(forced=1)
        self.nb.nl.check()

    def force(self, need_energy=0, need_virial=0, stream=None):
        self.nb.force_step(need_energy, need_virial, from_crd=True, overwrite=True, serial=self.serial,
                           stream=stream)

    def iterate(self, need_pressure=0, stream=None):
        md_iteration(self.settle, self.n, self.vel, self.nb.crd, self.nb.frc, self.acc, self.inverse_mass,
                     self.mass, self.nb.uint_crd, self.nb.box, self.dt, self.max_velocity, need_pressure,
                     self.pressure, stream)

    def volume(self):
        b = np.asarray(self.nb.box, np.float32)
        return float(np.float32(np.float32(b[0] * b[1]) * b[2]))

    def calculate_pressure(self, stream=None):
        """Calculate_Pressure_And_Potential_If_Needed's pressure
        (MD_core.cu:1765-1795) after a force step with need_virial: assigns
        self.pressure from the velocities and the step's FP64 parts."""
        pressure_from_parts(self.n, self.vel, self.mass, self.nb.parts, self.volume(), self.pressure, self.ek,
                            stream)

    def volume_change(self, factor, stream=None):
        """Main_Volume_Change (main.cu:438-448) after the caller scaled the
        coordinates: box *= factor (FP32, MD_core.cu:1490-1497), every module
        takes the new box, the list is checked (and rebuilt if needed) and the
        uint coordinates follow."""
        f = np.float32(factor)
        box = [float(f * np.float32(b)) for b in self.nb.box]
        self.nb.update_volume(box)
        self.nb.update_neighbor_list(forced=0, forced_check=1, stream=stream)
        self.nb.crd_to_uint(stream)

    def step(self, need_energy=0, stream=None, forced_update=None):
        """One MD step. forced_update None: the list's own interval / skin
        rule; 1: rebuild now; 0: never (used for graph capture). With a
        barostat the step also computes the pressure when the barostat asks
        for it, and scales the system after the integration (main.cu:334-372);
        such steps synchronise with the host."""
        baro = self.barostat
        need_pressure = bool(baro is not None and baro.Ask_For_Calculate_Pressure(self.steps + 1))
        self.force(need_energy, 1 if need_pressure else 0, stream)
        if need_pressure:
            self.calculate_pressure(stream)
        self.iterate(need_pressure=1 if need_pressure else 0, stream=stream)
        if need_pressure:
            import torch
            torch.cuda.synchronize()
            self.h_pressure = float(self.pressure.cpu()[0])
            if isinstance(baro, ANDERSEN_BAROSTAT_INFORMATION):
                mu, vs = baro.scale_factor(self.h_pressure, self.volume(), self.dt)
            else:
                mu, vs = baro.scale_factor(self.h_pressure)
            scale_list(self.nb.crd, mu, stream)
            if vs is not None:
                scale_list(self.vel, vs, stream)
            self.volume_change(mu, stream)
        if forced_update is None:
            self.nb.update_neighbor_list(stream=stream)
        elif forced_update:
            self.nb.update_neighbor_list(forced=1, stream=stream)
        self.steps += 1
        if need_energy or need_pressure:
            # steps that synchronise anyway report a list that left atoms or
            # pairs out (checked again after this step's rebuild, if any)
            import torch
            if not torch.cuda.is_current_stream_capturing():
                self.nb.nl.check()

    def degrees_of_freedom(self):
        """6 per rigid water, 3 per other atom."""
        return 6 * self.sys.waters + 3 * (self.n - 3 * self.sys.waters)

    def equilibrate(self, steps, temperature=300.0, stream=None):
        """Prepares a synthetic start: `steps` MD steps, each followed by a
        rescale of every velocity to `temperature` (the lattice start releases
        its strain energy as heat, which the rescale removes). This is input
        preparation, not a SPONGE thermostat; the rescale factor is computed
        on the device (no host round trip)."""
        import torch
        m = self.mass[:, None].double()
        target = 0.5 * self.degrees_of_freedom() * system.KB * float(temperature)
        for _ in range(int(steps)):
            self.step(stream=stream)
            ke = 0.5 * (m * self.vel.double() ** 2).sum()
            self.vel.mul_(torch.sqrt(target / torch.clamp(ke, min=1e-30)).float())
        torch.cuda.synchronize()

    def temperature(self):
        return 2.0 * self.kinetic_energy() / (self.degrees_of_freedom() * system.KB)

    def kinetic_energy(self):
        """Sum of m v^2 / 2 of the (half-step) velocities, FP64 on the host."""
        import torch
        torch.cuda.synchronize()
        v = self.vel.double().cpu().numpy()
        return 0.5 * float((self.sys.mass.astype(np.float64)[:, None] * v * v).sum())

    def potential_energy(self):
        p = self.nb.energy_parts()
        return p["lj"] + p["lrc"] + p["direct"] + p["correction"] + p["self"] + p["reciprocal"]


class _BatchMdIo(ctypes.Structure):
    """sponge_batch_md_io (include/sponge_md.h)."""
    _fields_ = [("vel", _P), ("acc", _P), ("inverse_mass", _P), ("mass", _P), ("dt", _F),
                ("max_velocity", _F)]


class BatchedMD:
    """MD steps of independent systems (FEP lambda windows, replicas of
    uneven size) together: each stage of `MDEngine.step` -- the fused force
    step, the leap-frog + SETTLE iteration and the neighbor-list update -- is
    ONE launch over every system (sponge_batch_force_step,
    sponge_batch_md_iteration, sponge_batch_neighbor_list_update), on the
    engines' own handles and device arrays. Per system the arithmetic is the
    batch's force step and sponge_md_iteration; the engines must share the
    list refresh interval. NVE only (no barostat, no pressure)."""

    def __init__(self, engines):
        self.engines = list(engines)
        k = len(self.engines)
        self.batch = BatchedSystems([e.nb for e in self.engines])
        settle = (_P * k)(*[e.settle.h if e.settle is not None else None for e in self.engines])
        io = (_BatchMdIo * k)()
        for i, e in enumerate(self.engines):
            io[i] = _BatchMdIo(_ptr(e.vel), _ptr(e.acc), _ptr(e.inverse_mass), _ptr(e.mass), float(e.dt),
                               float(e.max_velocity))
        self._keep = (settle, io)
        _check(load_library().sponge_batch_set_integrator(self.batch.h, ctypes.cast(settle, _P),
                                                          ctypes.cast(io, _P)), "sponge_batch_set_integrator")
        self.refresh_interval = self.engines[0].nb.refresh_interval if self.engines else 0
        self.steps = 0

    def prepare(self):
        """Main_Initial's forced list build of every system; raises if a list
        overflowed a capacity."""
        self.batch.neighbor_list_update(forced=1)
        for e in self.engines:
            e.nb.nl.check()

    def iterate(self, stream=None):
        _check(load_library().sponge_batch_md_iteration(self.batch.h, _stream(stream)),
               "sponge_batch_md_iteration")

    def step(self, stream=None, forced_update=None):
        """One MD step of every system. forced_update None: the lists'
        interval rule (host counter, as sponge_batch_neighbor_list_update);
        1: rebuild now; 0: no list update (graph capture of a plain step)."""
        self._launch(stream, forced_update)
        self.steps += 1

    def _launch(self, stream, forced_update):
        self.batch.force_step(0, 0, from_crd=True, overwrite=True, stream=stream)
        self.iterate(stream)
        if forced_update is None:
            self.batch.neighbor_list_update(stream=stream)
        elif forced_update:
            self.batch.neighbor_list_update(forced=1, stream=stream)

    def capture(self):
        """One hipGraph per MD step form: (plain step, step with the list
        rebuild). Replay the second on the steps where `step()` would
        rebuild (every refresh_interval-th update, the first included)."""
        import torch
        plain, rebuild = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(plain):
            self._launch(torch.cuda.current_stream(), 0)
        with torch.cuda.graph(rebuild):
            self._launch(torch.cuda.current_stream(), 1)
        return plain, rebuild
