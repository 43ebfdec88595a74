"""Python mirror of SPONGE's three nonbonded modules over the C ABI.

`NEIGHBOR_LIST`, `LENNARD_JONES_INFORMATION` and `Particle_Mesh_Ewald` keep
the reference's method names and argument meaning
(neighbor_list/neighbor_list.cuh:83-105, Lennard_Jones_force/
Lennard_Jones_force.cuh:72-101, PME_force/PME_force.cuh:96-121); every call
goes straight to libsponge_nb.so (hand-written HIP for gfx950). Device memory
and streams come from PyTorch; there is no CPU fallback: without the built
library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsponge_nb.so")
_lib = None

NPARTS = 8
OVERWRITE_FORCE = 1
SERIAL = 2
FROM_CRD = 4
TWO_STREAMS = 8
# sponge_nl_status overflow bits
OVERFLOW_COLUMN = 1
OVERFLOW_LIST = 2
OVERFLOW_CANDIDATES = 4
PART_NAMES = ["lj", "direct", "lj_virial", "correction", "self", "reciprocal", "lrc", "pairs_in"]

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_L = ctypes.c_long
_D = ctypes.c_double
_FP3 = ctypes.POINTER(ctypes.c_float)

_SIGS = {
    "sponge_nb_last_error": (ctypes.c_char_p, []),
    "sponge_nb_device_info": (_I, [ctypes.POINTER(_I), ctypes.c_char_p, _I]),
    "sponge_crd_to_uint": (_I, [_I, _FP3, _P, _P, _P]),
    "sponge_nl_create": (_I, [_I, _FP3, _F, _F, _I, _I, _I, _F, ctypes.POINTER(_P)]),
    "sponge_nl_destroy": (None, [_P]),
    "sponge_nl_update": (_I, [_P, _P, _P, _P, _P, _I, _I, _P]),
    "sponge_nl_update_volume": (_I, [_P, _FP3]),
    "sponge_nl_set_prune": (_I, [_P, _I, _F]),
    "sponge_nl_inner_status": (_I, [_P, _P, _P]),
    "sponge_nl_prune_votes": (_I, [_P, ctypes.POINTER(_I)]),
    "sponge_nl_status": (_I, [_P, ctypes.POINTER(_L), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "sponge_nl_set_exclusions": (_I, [_P, _P, _P, _P]),
    "sponge_nl_get_pairs": (_I, [_P, _P, _L, ctypes.POINTER(_L)]),
    "sponge_lj_create": (_I, [_I, _I, _P, _P, _P, _F, _FP3, ctypes.POINTER(_P)]),
    "sponge_lj_destroy": (None, [_P]),
    "sponge_lj_pme_direct_force": (_I, [_P, _P, _P, _P, _P, _F, _I, _P, _I, _P, _P, _P]),
    "sponge_lj_pme_direct_force_noalloc": (_I, [_P, _P, _P, _P, _P, _F, _I, _P, _I, _P, _P, _P]),
    "sponge_lj_force_with_pme_direct_force": (_I, [_P, _P, _P, _P, _P, _F, _P]),
    "sponge_lj_force_with_pme_direct_force_records": (_I, [_P, _P, _I, _P, _FP3, _P, _F, _F, _P]),
    "sponge_lj_long_range_correction": (_I, [_P, _I, _P, _I, _P, _P]),
    "sponge_lj_get_energy": (_I, [_P, _P, _P, ctypes.POINTER(_D), _P]),
    "sponge_lj_update_volume": (_I, [_P, _FP3]),
    "sponge_lj_host_tables": (_I, [_P, _P, _P, _P]),
    "sponge_lj_set_pair_events": (_I, [_P, _P, _P]),
    "sponge_lj_soft_create": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _F, _F, _F, _F, _F, _F, _FP3,
                                   ctypes.POINTER(_P)]),
    "sponge_pme_create": (_I, [_I, _FP3, _F, _F, _I, _I, _I, ctypes.POINTER(_P)]),
    "sponge_pme_destroy": (None, [_P]),
    "sponge_pme_info": (_I, [_P, ctypes.POINTER(_F), ctypes.POINTER(_I)]),
    "sponge_pme_excluded_force": (_I, [_P, _P, _FP3, _P, _P, _P, _P, _P, _P, _P]),
    "sponge_pme_reciprocal_force": (_I, [_P, _P, _P, _P, _I, _I, _P, _P, _P, _P]),
    "sponge_pme_get_energy": (_I, [_P, _P, _P, _P, _FP3, _P, _P, _P, _I, ctypes.POINTER(_D), _P]),
    "sponge_pme_update_box_length": (_I, [_P, _FP3]),
    "sponge_pme_set_events": (_I, [_P, _P, _P, _P, _P]),
    "sponge_pme_set_tuning": (_I, [_P, _I, _I]),
    "sponge_nb_default_shapes": (_I, [_I, _I, _F, _F, _P]),
    "sponge_nb_get_shapes": (_I, [_P, _P, _P, _P]),
    "sponge_nb_set_shapes": (_I, [_P, _P, _P, _P]),
    "sponge_nb_force_step": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "sponge_sum_list": (_I, [_I, _P, _P, _P]),
    "sponge_lj_set_tables": (_I, [_P, _P, _P]),
    "sponge_nb_energy_gradient": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sponge_nb_force_matching_gradient": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sponge_batch_create": (_I, [_I, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "sponge_batch_destroy": (None, [_P]),
    "sponge_batch_sync": (_I, [_P]),
    "sponge_batch_neighbor_list_update": (_I, [_P, _I, _P]),
    "sponge_batch_force_step": (_I, [_P, _I, _I, _I, _P]),
    "sponge_batch_set_integrator": (_I, [_P, _P, _P]),
    "sponge_batch_md_iteration": (_I, [_P, _P]),
    # include/sponge_md.h
    "sponge_md_leap_frog": (_I, [_I, _P, _P, _P, _P, _P, _F, _F, _P]),
    "sponge_settle_create": (_I, [_I, _I, _P, _P, _P, _P, _F, _F, _I, ctypes.POINTER(_P)]),
    "sponge_settle_destroy": (None, [_P]),
    "sponge_settle_counts": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "sponge_settle_groups": (_I, [_P, _P, _P, _P, _P]),
    "sponge_settle_set_factors": (_I, [_P, _F, _F]),
    "sponge_settle_remember_last_coordinates": (_I, [_P, _P, _FP3, _P]),
    "sponge_settle_do": (_I, [_P, _P, _P, _FP3, _P, _I, _P, _P]),
    "sponge_md_iteration": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _FP3, _F, _F, _I, _P, _P]),
    "sponge_md_get_pressure": (_I, [_I, _P, _P, _P, _P, _F, _P, _P, _P]),
    "sponge_md_pressure_from_parts": (_I, [_I, _P, _P, _P, _F, _P, _P, _P]),
    "sponge_md_scale_list": (_I, [_I, _P, _F, _P]),
    # include/sponge_fep.h
    "sponge_lj_soft_energy_with_coulomb_direct": (_I, [_P, _P, _P, _P, _F, _P, _P]),
    "sponge_lj_soft_dhdl_with_coulomb_direct": (_I, [_P, _P, _P, _P, _P, _I, _F, _P, _P]),
    "sponge_lj_soft_long_range_correction": (_I, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "sponge_lj_soft_lambda_windows": (_I, [_P, _P, _P, _P, _P, _F, _I, _P, _P, _P]),
    "sponge_pme_cross_dhdl": (_I, [_P, _P, _P, _P, _P, _FP3, _P, _P, _P, _I, _P, _P]),
    "sponge_pme_lambda_windows": (_I, [_P, _P, _P, _P, _FP3, _P, _P, _P, _F, _I, _P, _P, _P]),
    "sponge_fep_separate_direct_energy": (_I, [_P, _P, _P, _P, _P, _P, _P]),
}

FEP_FIELDS = ["vdw_intersys", "vdw_intrasys", "coul_direct_intersys", "coul_direct_intrasys", "dhdl_lj",
              "dhdl_direct", "dhdl_direct_exact", "vdw_lrc", "dhdl_lrc"]


class SpongeError(RuntimeError):
    pass


class Shapes(ctypes.Structure):
    """sponge_shapes (include/sponge_nb.h): launch shapes and schedule of the
    fused step."""
    _fields_ = [("pair_waves", _I), ("conv_split", _I), ("conv_threads", _I), ("step_form", _I),
                ("prune_interval", _I), ("prune_radius", _F)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["prune_radius"] = round(float(d["prune_radius"]), 4)
        return d

    @classmethod
    def from_dict(cls, d):
        return cls(**{name: d[name] for name, _ in cls._fields_})


def load_library(path: str = LIB_PATH):
    """Loads libsponge_nb.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SpongeError(f"{path} is missing: run `python -m sponge_amd.build` first "
                          "(there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


def _check(rc: int, what: str):
    if rc != 0:
        msg = load_library().sponge_nb_last_error().decode(errors="replace")
        raise SpongeError(f"{what} failed ({rc}): {msg}")


def _vec3(v):
    a = (ctypes.c_float * 3)(*[float(x) for x in v])
    return a


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise SpongeError("no GPU visible: the HIP path has no CPU fallback")
    lib = load_library()
    cnt = _I(0)
    name = ctypes.create_string_buffer(64)
    torch.cuda.init()
    rc = lib.sponge_nb_device_info(ctypes.byref(cnt), name, 64)
    _check(rc, "device check")
    return name.value.decode()


class NEIGHBOR_LIST:
    """neighbor_list/neighbor_list.cuh:44-105 over sponge_nl_*."""
    CONDITIONAL_UPDATE, FORCED_UPDATE = 0, 1
    CONDITIONAL_CHECK, FORCED_CHECK = 0, 1

    def __init__(self, atom_numbers, box_length, cut, skin, refresh_interval=0,
                 max_atom_in_grid_numbers=64, max_neighbor_numbers=800, skin_permit=0.5):
        lib = load_library()
        h = _P()
        _check(lib.sponge_nl_create(int(atom_numbers), _vec3(box_length), float(cut), float(skin),
                                    int(refresh_interval), int(max_atom_in_grid_numbers),
                                    int(max_neighbor_numbers), float(skin_permit), ctypes.byref(h)),
               "NEIGHBOR_LIST::Initial")
        self.h = h
        self.atom_numbers = int(atom_numbers)

    def Neighbor_List_Update(self, crd, excluded_list_start, excluded_list, excluded_numbers,
                             forced_update=0, forced_check=0, stream=None):
        _check(load_library().sponge_nl_update(self.h, _ptr(crd), _ptr(excluded_list_start),
                                               _ptr(excluded_list), _ptr(excluded_numbers),
                                               int(forced_update), int(forced_check), _stream(stream)),
               "Neighbor_List_Update")

    def set_exclusions(self, excluded_list_start, excluded_list, excluded_numbers):
        """Rebuilds the list's symmetric copy of the exclusion CSR (after the
        caller rewrote the arrays in place). Synchronises."""
        _check(load_library().sponge_nl_set_exclusions(self.h, _ptr(excluded_list_start), _ptr(excluded_list),
                                                       _ptr(excluded_numbers)), "sponge_nl_set_exclusions")

    def set_prune(self, interval, radius=0.0):
        """Inner list of the fused step: per-atom lists of the full list's
        pairs, written by each rebuild, pruned every
        `interval` fused steps to `radius` (>= cutoff) when `radius` is below
        cutoff + skin; interval 0 turns it off."""
        _check(load_library().sponge_nl_set_prune(self.h, int(interval), float(radius)), "sponge_nl_set_prune")

    def Update_Volume(self, box_length):
        _check(load_library().sponge_nl_update_volume(self.h, _vec3(box_length)), "Update_Volume")

    def status(self):
        pc, fl, rb = _L(0), _I(0), _I(0)
        rc = load_library().sponge_nl_status(self.h, ctypes.byref(pc), ctypes.byref(fl), ctypes.byref(rb))
        return dict(pairs=pc.value, overflow=fl.value, rebuilds=rb.value, ok=rc == 0)

    def check(self):
        """Raises SpongeError when the last rebuild overflowed a capacity
        (SPONGE_NL_OVERFLOW_*: atoms or pairs were left out of the list).
        Synchronises with the device: call it between steps, outside a graph
        capture."""
        st = self.status()
        if not st["ok"]:
            raise SpongeError(f"neighbor list capacity overflow (flags {st['overflow']:#x}): "
                              "atoms or pairs were left out of the last rebuild")
        return st

    def inner_status(self):
        """Per-slot inner lists of the fused step: overflow at the last prune
        and the number of pair-kernel launches that read them."""
        ov, nl_ = _I(0), _I(0)
        _check(load_library().sponge_nl_inner_status(self.h, ctypes.byref(ov), ctypes.byref(nl_)),
               "sponge_nl_inner_status")
        return dict(slot_overflow=ov.value, slot_launches=nl_.value)

    def prune_votes(self):
        """Inner-list prunes caused by the displacement test (beyond the
        scheduled ones) since the list was created."""
        v = _I(0)
        _check(load_library().sponge_nl_prune_votes(self.h, ctypes.byref(v)), "sponge_nl_prune_votes")
        return v.value

    def pairs(self):
        """(P,2) int32 canonical pairs (i<j), sorted lexicographically."""
        lib = load_library()
        n = _L(0)
        lib.sponge_nl_get_pairs(self.h, None, 0, ctypes.byref(n))
        out = np.zeros((max(n.value, 1), 2), np.int32)
        rc = lib.sponge_nl_get_pairs(self.h, out.ctypes.data_as(_P), n.value, ctypes.byref(n))
        _check(rc, "neighbor list capacity")
        out = out[:n.value]
        order = np.lexsort((out[:, 1], out[:, 0]))
        return out[order]

    def Clear(self):
        if self.h:
            load_library().sponge_nl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.Clear()
        except Exception:
            pass


class LENNARD_JONES_INFORMATION:
    """Lennard_Jones_force/Lennard_Jones_force.cuh:37-149 over sponge_lj_*."""

    def __init__(self, atom_numbers, atom_type_numbers, h_LJ_A, h_LJ_B, h_atom_LJ_type, cutoff,
                 box_length):
        lib = load_library()
        self._A = np.ascontiguousarray(h_LJ_A, np.float32)
        self._B = np.ascontiguousarray(h_LJ_B, np.float32)
        self._t = np.ascontiguousarray(h_atom_LJ_type, np.int32)
        h = _P()
        _check(lib.sponge_lj_create(int(atom_numbers), int(atom_type_numbers), self._A.ctypes.data_as(_P),
                                    self._B.ctypes.data_as(_P), self._t.ctypes.data_as(_P), float(cutoff),
                                    _vec3(box_length), ctypes.byref(h)),
               "LENNARD_JONES_INFORMATION::Initial")
        self.h = h
        self.cutoff = float(cutoff)

    def LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(self, nl, uint_crd, charge, frc, pme_beta,
                                                         need_atom_energy=0, atom_energy=None,
                                                         need_virial=0, atom_lj_virial=None,
                                                         atom_direct_pme_energy=None, stream=None):
        _check(load_library().sponge_lj_pme_direct_force(
            self.h, nl.h, _ptr(uint_crd), _ptr(charge), _ptr(frc), float(pme_beta), int(need_atom_energy),
            _ptr(atom_energy), int(need_virial), _ptr(atom_lj_virial), _ptr(atom_direct_pme_energy),
            _stream(stream)), "LJ_PME_Direct_Force_With_Atom_Energy_And_Virial")

    def LJ_Force_With_PME_Direct_Force(self, *args, stream=None):
        """Lennard_Jones_force.cuh:121-124, both overloads:
        (uint_crd, charge, frc, nl, pme_beta) or
        (atom_numbers, uint_crd_with_LJ, scaler, frc, nl, cutoff, pme_beta),
        the records an (N, 5) int32 device tensor (uint x, y, z, LJ type,
        charge bits)."""
        lib = load_library()
        if len(args) == 5:
            uint_crd, charge, frc, nl, pme_beta = args
            _check(lib.sponge_lj_force_with_pme_direct_force(self.h, nl.h, _ptr(uint_crd), _ptr(charge), _ptr(frc),
                                                            float(pme_beta), _stream(stream)),
                   "LJ_Force_With_PME_Direct_Force")
        elif len(args) == 7:
            atom_numbers, records, scaler, frc, nl, cutoff, pme_beta = args
            _check(lib.sponge_lj_force_with_pme_direct_force_records(self.h, nl.h, int(atom_numbers), _ptr(records),
                                                                    _vec3(scaler), _ptr(frc), float(cutoff),
                                                                    float(pme_beta), _stream(stream)),
                   "LJ_Force_With_PME_Direct_Force")
        else:
            raise TypeError("LJ_Force_With_PME_Direct_Force takes 5 or 7 arguments")

    def Long_Range_Correction(self, need_pressure, d_virial, need_potential, d_potential, stream=None):
        _check(load_library().sponge_lj_long_range_correction(
            self.h, int(need_pressure), _ptr(d_virial), int(need_potential), _ptr(d_potential),
            _stream(stream)), "Long_Range_Correction")

    def Get_Energy(self, nl, uint_crd, stream=None):
        e = _D(0)
        _check(load_library().sponge_lj_get_energy(self.h, nl.h, _ptr(uint_crd), ctypes.byref(e),
                                                   _stream(stream)), "LJ Get_Energy")
        return e.value

    def Update_Volume(self, box_length):
        _check(load_library().sponge_lj_update_volume(self.h, _vec3(box_length)), "LJ Update_Volume")

    def set_tables(self, h_LJ_A, h_LJ_B):
        """New x12 / x6 triangular tables (plain handles)."""
        self._A = np.ascontiguousarray(h_LJ_A, np.float32)
        self._B = np.ascontiguousarray(h_LJ_B, np.float32)
        _check(load_library().sponge_lj_set_tables(self.h, self._A.ctypes.data_as(_P), self._B.ctypes.data_as(_P)),
               "sponge_lj_set_tables")

    def set_pair_events(self, start=None, stop=None):
        """Brackets later pair-kernel launches with these torch.cuda.Events."""
        e0 = ctypes.c_void_p(start.cuda_event) if start is not None else None
        e1 = ctypes.c_void_p(stop.cuda_event) if stop is not None else None
        _check(load_library().sponge_lj_set_pair_events(self.h, e0, e1), "set_pair_events")

    def Clear(self):
        if self.h:
            load_library().sponge_lj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.Clear()
        except Exception:
            pass


class LJ_SOFT_CORE(LENNARD_JONES_INFORMATION):
    """Lennard_Jones_force/LJ_soft_core.cuh:37-160 over sponge_lj_soft_create;
    the force, long-range, energy and volume methods are those of the plain
    module (the handle type is shared). Tables are x12 / x6 as stored."""

    def __init__(self, atom_numbers, atom_type_numbers_A, atom_type_numbers_B, h_LJ_AA, h_LJ_AB, h_LJ_BA,
                 h_LJ_BB, h_atom_LJ_type_A, h_atom_LJ_type_B, h_subsys_division, lambda_lj, cutoff,
                 box_length, soft_core_alpha=0.5, soft_core_powfer=1.0, soft_core_sigma=3.0,
                 soft_core_sigma_min=0.0):
        lib = load_library()
        keep = [np.ascontiguousarray(a, dt) for a, dt in (
            (h_LJ_AA, np.float32), (h_LJ_AB, np.float32), (h_LJ_BA, np.float32), (h_LJ_BB, np.float32),
            (h_atom_LJ_type_A, np.int32), (h_atom_LJ_type_B, np.int32))]
        mask = None if h_subsys_division is None else np.ascontiguousarray(h_subsys_division, np.int32)
        self._keep = keep + [mask]
        h = _P()
        _check(lib.sponge_lj_soft_create(
            int(atom_numbers), int(atom_type_numbers_A), int(atom_type_numbers_B),
            *[k.ctypes.data_as(_P) for k in keep], mask.ctypes.data_as(_P) if mask is not None else None,
            float(lambda_lj), float(soft_core_alpha), float(soft_core_powfer), float(soft_core_sigma),
            float(soft_core_sigma_min), float(cutoff), _vec3(box_length), ctypes.byref(h)),
            "LJ_SOFT_CORE::Initial")
        self.h = h
        self.cutoff = float(cutoff)
        self.lambda_lj = float(lambda_lj)
        self.h_LJ_energy_sum_intersys = self.h_LJ_energy_sum_intrasys = 0.0
        self.h_direct_ene_sum_intersys = self.h_direct_ene_sum_intrasys = 0.0
        self.h_sigma_of_dH_dlambda_lj = self.h_sigma_of_dH_dlambda_direct = 0.0

    def Get_Energy_With_Coulomb_Direct(self, uint_crd, nl, charge, pme_beta, stream=None):
        """LJ_soft_core.cu:1809-1852: fills the four h_*_sum_* members and
        returns the LJ total (intersys + intrasys)."""
        out = np.zeros(4)
        _check(load_library().sponge_lj_soft_energy_with_coulomb_direct(
            self.h, nl.h, _ptr(uint_crd), _ptr(charge), float(pme_beta), out.ctypes.data_as(_P),
            _stream(stream)), "Get_Energy_With_Coulomb_Direct")
        (self.h_LJ_energy_sum_intersys, self.h_LJ_energy_sum_intrasys, self.h_direct_ene_sum_intersys,
         self.h_direct_ene_sum_intrasys) = (float(x) for x in out)
        return self.h_LJ_energy_sum_intersys + self.h_LJ_energy_sum_intrasys

    def Get_Partial_H_Partial_Lambda_With_Columb_Direct(self, uint_crd, charge, nl, charge_B_A,
                                                         charge_pertubated, pme_beta, stream=None):
        """LJ_soft_core.cu:1854-1894: returns the LJ dH/dlambda and keeps the
        direct-Coulomb one in h_sigma_of_dH_dlambda_direct."""
        out = np.zeros(2)
        _check(load_library().sponge_lj_soft_dhdl_with_coulomb_direct(
            self.h, nl.h, _ptr(uint_crd), _ptr(charge), _ptr(charge_B_A), int(charge_pertubated),
            float(pme_beta), out.ctypes.data_as(_P), _stream(stream)),
            "Get_Partial_H_Partial_Lambda_With_Columb_Direct")
        self.h_sigma_of_dH_dlambda_lj, self.h_sigma_of_dH_dlambda_direct = float(out[0]), float(out[1])
        return self.h_sigma_of_dH_dlambda_lj

    def Long_Range_Correction(self, *args, **kw):
        """No arguments: the host value factor / V (LJ_soft_core.cu:1927-1937);
        with arguments: the device accumulation of the plain module."""
        if args or kw:
            return super().Long_Range_Correction(*args, **kw)
        v = _D(0)
        _check(load_library().sponge_lj_soft_long_range_correction(self.h, ctypes.byref(v), None),
               "Long_Range_Correction")
        return v.value

    def Partial_H_Partial_Lambda_Long_Range_Correction(self):
        """LJ_soft_core.cu:1939-1955."""
        v = _D(0)
        _check(load_library().sponge_lj_soft_long_range_correction(self.h, None, ctypes.byref(v)),
               "Partial_H_Partial_Lambda_Long_Range_Correction")
        return v.value

    def lambda_windows(self, nl, uint_crd, charge, charge_B_A, pme_beta, lambdas, stream=None):
        """Every window of one frame (sponge_lj_soft_lambda_windows): array
        (nwin, 9) in FEP_FIELDS order."""
        lams = np.ascontiguousarray(np.atleast_1d(lambdas), np.float64)
        out = np.zeros((len(lams), len(FEP_FIELDS)))
        _check(load_library().sponge_lj_soft_lambda_windows(
            self.h, nl.h, _ptr(uint_crd), _ptr(charge), _ptr(charge_B_A), float(pme_beta), len(lams),
            lams.ctypes.data_as(_P), out.ctypes.data_as(_P), _stream(stream)), "lambda windows")
        return out


class Particle_Mesh_Ewald:
    """PME_force/PME_force.cuh:25-144 over sponge_pme_*."""
    TOTAL, DIRECT, RECIPROCAL, CORRECTION, SELF = 0, 1, 2, 3, 4

    def __init__(self, atom_numbers, boxlength, cutoff, direct_tolerance=1e-5, fftx=-1, ffty=-1,
                 fftz=-1):
        lib = load_library()
        h = _P()
        _check(lib.sponge_pme_create(int(atom_numbers), _vec3(boxlength), float(cutoff),
                                     float(direct_tolerance), int(fftx), int(ffty), int(fftz),
                                     ctypes.byref(h)), "Particle_Mesh_Ewald::Initial")
        self.h = h
        b = _F(0)
        k = (_I * 3)()
        lib.sponge_pme_info(h, ctypes.byref(b), k)
        self.beta = b.value
        self.fftx, self.ffty, self.fftz = k[0], k[1], k[2]

    def PME_Excluded_Force_With_Atom_Energy(self, uint_crd, scaler, charge, excluded_list_start,
                                            excluded_list, excluded_atom_numbers, frc, atom_energy=None,
                                            stream=None):
        _check(load_library().sponge_pme_excluded_force(
            self.h, _ptr(uint_crd), _vec3(scaler), _ptr(charge), _ptr(excluded_list_start),
            _ptr(excluded_list), _ptr(excluded_atom_numbers), _ptr(frc), _ptr(atom_energy),
            _stream(stream)), "PME_Excluded_Force_With_Atom_Energy")

    def PME_Reciprocal_Force_With_Energy_And_Virial(self, uint_crd, charge, force, need_virial=0,
                                                    need_energy=0, d_virial=None, d_potential=None,
                                                    d_direct_atom_energy=None, stream=None):
        _check(load_library().sponge_pme_reciprocal_force(
            self.h, _ptr(uint_crd), _ptr(charge), _ptr(force), int(need_virial), int(need_energy),
            _ptr(d_virial), _ptr(d_potential), _ptr(d_direct_atom_energy), _stream(stream)),
            "PME_Reciprocal_Force_With_Energy_And_Virial")

    def Get_Energy(self, uint_crd, charge, nl, scaler, excluded_list_start, excluded_list,
                   excluded_atom_numbers, which_part=0, stream=None):
        e = _D(0)
        _check(load_library().sponge_pme_get_energy(
            self.h, nl.h if nl is not None else None, _ptr(uint_crd), _ptr(charge), _vec3(scaler),
            _ptr(excluded_list_start), _ptr(excluded_list), _ptr(excluded_atom_numbers), int(which_part),
            ctypes.byref(e), _stream(stream)), "PME Get_Energy")
        return e.value

    def cross_dhdl(self, nl, uint_crd, charge, charge_B_A, scaler, excluded_list_start, excluded_list,
                   excluded_atom_numbers, lj_pertubated, stream=None):
        """TI_CORE::Get_Cross_PME_Partial_H_Partial_Lambda (TI_core.cu:670-739):
        dict of the reciprocal, self, correction and direct cross terms."""
        out = np.zeros(4)
        _check(load_library().sponge_pme_cross_dhdl(
            self.h, nl.h if nl is not None else None, _ptr(uint_crd), _ptr(charge), _ptr(charge_B_A),
            _vec3(scaler), _ptr(excluded_list_start), _ptr(excluded_list), _ptr(excluded_atom_numbers),
            int(lj_pertubated), out.ctypes.data_as(_P), _stream(stream)), "Get_Cross_PME_Partial_H_Partial_Lambda")
        return dict(reciprocal=float(out[0]), self=float(out[1]), correction=float(out[2]), direct=float(out[3]))

    def lambda_windows(self, uint_crd, charge, charge_B_A, scaler, excluded_list_start, excluded_list,
                       excluded_atom_numbers, lambda0, lambdas, stream=None):
        """Reciprocal, self and correction energies at every window
        (sponge_pme_lambda_windows): array (nwin, 3)."""
        lams = np.ascontiguousarray(np.atleast_1d(lambdas), np.float64)
        out = np.zeros((len(lams), 3))
        _check(load_library().sponge_pme_lambda_windows(
            self.h, _ptr(uint_crd), _ptr(charge), _ptr(charge_B_A), _vec3(scaler), _ptr(excluded_list_start),
            _ptr(excluded_list), _ptr(excluded_atom_numbers), float(lambda0), len(lams), lams.ctypes.data_as(_P),
            out.ctypes.data_as(_P), _stream(stream)), "PME lambda windows")
        return out

    def Seperate_Direct_Atom_Energy(self, nl, uint_crd, charge, subsys_division=None, stream=None):
        """FEP_CORE::Seperate_Direct_Atom_Energy (FEP_core.cu:545-558) with this
        module's cutoff and beta: (intersys, intrasys)."""
        out = np.zeros(2)
        _check(load_library().sponge_fep_separate_direct_energy(
            self.h, nl.h, _ptr(uint_crd), _ptr(charge), _ptr(subsys_division), out.ctypes.data_as(_P),
            _stream(stream)), "Seperate_Direct_Atom_Energy")
        return float(out[0]), float(out[1])

    def set_tuning(self, step_form=-1, conv_split=-1):
        """Step form (0: pair kernel on a second stream, 1: every kernel on
        its own) and convolution form (0: one workgroup per plane, 1: line
        launches); -1 keeps a setting."""
        _check(load_library().sponge_pme_set_tuning(self.h, int(step_form), int(conv_split)),
               "sponge_pme_set_tuning")

    def set_events(self, spread=None, gather=None):
        """Measurement hook: (start, stop) torch.cuda.Event pairs recorded by
        the spread and gather launches themselves; None switches one off."""
        def ev(e):
            return None if e is None else ctypes.c_void_p(e.cuda_event)
        sp = spread or (None, None)
        ga = gather or (None, None)
        _check(load_library().sponge_pme_set_events(self.h, ev(sp[0]), ev(sp[1]), ev(ga[0]), ev(ga[1])),
               "sponge_pme_set_events")

    def Update_Box_Length(self, boxlength):
        _check(load_library().sponge_pme_update_box_length(self.h, _vec3(boxlength)),
               "PME Update_Box_Length")

    Update_Volume = Update_Box_Length

    def Clear(self):
        if self.h:
            load_library().sponge_pme_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.Clear()
        except Exception:
            pass


def crd_to_uint(crd, box_length, uint_out, stream=None):
    """MD_INFORMATION::MD_Information_Crd_To_Uint_Crd (MD_core/MD_core.cu:1528)."""
    _check(load_library().sponge_crd_to_uint(int(crd.shape[0]), _vec3(box_length), _ptr(crd),
                                              _ptr(uint_out), _stream(stream)), "Crd_To_Uint_Crd")


def sum_list(lst, out, stream=None):
    _check(load_library().sponge_sum_list(int(lst.numel()), _ptr(lst), _ptr(out), _stream(stream)),
           "Sum_Of_List")


class NonbondedEngine:
    """The three modules plus the device state MD_INFORMATION owns for them.

    Built from a `sponge_amd.system.System`; `force_step` runs the fused
    nonbonded force of one MD step (main.cu:180-193) into `self.frc`.
    """

    def __init__(self, sys_, refresh_interval=20, device="cuda"):
        import torch
        require_gpu()
        self.sys = sys_
        self.device = device
        n = sys_.n
        self.n = n
        self.box = [float(x) for x in sys_.box]
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
        self.crd = t(sys_.crd, torch.float32).contiguous()
        self.charge = t(sys_.charge, torch.float32)
        self.uint_crd = torch.zeros((n, 3), dtype=torch.int32, device=device)
        self.frc = torch.zeros((n, 3), dtype=torch.float32, device=device)
        self.atom_energy = torch.zeros(n, dtype=torch.float32, device=device)
        self.atom_virial = torch.zeros(n, dtype=torch.float32, device=device)
        self.atom_direct = torch.zeros(n, dtype=torch.float32, device=device)
        self.parts = torch.zeros(NPARTS, dtype=torch.float64, device=device)
        self.excl_start = t(sys_.excl_start, torch.int32)
        self.excl_numbers = t(sys_.excl_numbers, torch.int32)
        el = sys_.excl_list if len(sys_.excl_list) else np.zeros(1, np.int32)
        self.excl_list = t(el, torch.int32)
        self.nl = NEIGHBOR_LIST(n, self.box, sys_.cutoff, sys_.skin, refresh_interval=refresh_interval)
        self.refresh_interval = int(refresh_interval)
        self._interval_updates = 0
        if getattr(sys_, "lj_type_B", None) is not None:
            self.lj = LJ_SOFT_CORE(n, sys_.ntypes, sys_.ntypes_B, sys_.lj_A, sys_.lj_B, sys_.lj_A_B,
                                   sys_.lj_B_B, sys_.lj_type, sys_.lj_type_B, sys_.mask, sys_.lam,
                                   sys_.cutoff, self.box, sys_.sc_alpha, sys_.sc_p, sys_.sc_sigma,
                                   sys_.sc_sigma_min)
        else:
            self.lj = LENNARD_JONES_INFORMATION(n, sys_.ntypes, sys_.lj_A, sys_.lj_B, sys_.lj_type,
                                                sys_.cutoff, self.box)
        self.pme = Particle_Mesh_Ewald(n, self.box, sys_.cutoff, sys_.pme_tolerance, *sys_.fft)
        self.scaler = [b / 4294967296.0 for b in self.box]

    def update_neighbor_list(self, forced=0, stream=None, forced_check=0):
        self.nl.Neighbor_List_Update(self.crd, self.excl_start, self.excl_list, self.excl_numbers,
                                     forced_update=forced, forced_check=forced_check, stream=stream)
        if not forced and not forced_check and self.refresh_interval > 0:
            self._interval_updates += 1

    def steps_to_refresh(self):
        """Interval mode: force steps left before the update that rebuilds
        the list (1 = the next force step is the last on this list)."""
        r = self.refresh_interval
        return (r - self._interval_updates % r) % r + 1

    def update_volume(self, box):
        """The module half of Main_Volume_Change / Main_Box_Length_Change
        (main.cu:438-461): every module takes the new box (the PME module its
        new box, not the stale one the reference's Update_Volume forwards,
        PME_force.cu:996-999). Host-side: issue between steps, not inside a
        captured graph."""
        self.box = [float(x) for x in box]
        self.nl.Update_Volume(self.box)
        self.lj.Update_Volume(self.box)
        self.pme.Update_Box_Length(self.box)
        self.scaler = [b / 4294967296.0 for b in self.box]

    def crd_to_uint(self, stream=None):
        crd_to_uint(self.crd, self.box, self.uint_crd, stream)

    def force_step(self, need_energy=0, need_virial=0, from_crd=False, overwrite=False, serial=False,
                   stream=None, two_streams=False):
        """Fused nonbonded force of one step into self.frc. from_crd: compute the
        uint coordinates from self.crd first; overwrite: assign self.frc
        instead of adding to it; serial: every kernel on its own, one after
        the other; two_streams: the pair kernel on a second stream; neither:
        the PME handle's tuned form."""
        flags = ((OVERWRITE_FORCE if overwrite else 0) | (SERIAL if serial else 0) |
                 (TWO_STREAMS if two_streams else 0))
        _check(load_library().sponge_nb_force_step(
            self.nl.h, self.lj.h, self.pme.h, _ptr(self.crd) if from_crd else None, _ptr(self.uint_crd),
            _ptr(self.charge), _ptr(self.excl_start), _ptr(self.excl_list), _ptr(self.excl_numbers),
            _ptr(self.frc), int(flags), int(need_energy), int(need_virial), _ptr(self.atom_energy),
            _ptr(self.atom_virial), _ptr(self.parts), _stream(stream)), "nb_force_step")

    def energy_gradient(self, grad_charge, grad_A, grad_B, stream=None):
        """dE/dq, dE/dA, dE/dB (float64 device tensors) of the total energy
        at the current coordinates (the list must be current)."""
        _check(load_library().sponge_nb_energy_gradient(
            self.nl.h, self.lj.h, self.pme.h, _ptr(self.uint_crd), _ptr(self.charge), _ptr(self.excl_start),
            _ptr(self.excl_list), _ptr(self.excl_numbers), _ptr(grad_charge), _ptr(grad_A), _ptr(grad_B),
            _stream(stream)), "sponge_nb_energy_gradient")

    def force_matching_gradient(self, frc, target, loss, grad_charge, grad_A, grad_B, stream=None):
        """L = sum |frc - target|^2 into loss[0] and its gradients."""
        _check(load_library().sponge_nb_force_matching_gradient(
            self.nl.h, self.lj.h, self.pme.h, _ptr(self.uint_crd), _ptr(self.charge), _ptr(self.excl_start),
            _ptr(self.excl_list), _ptr(self.excl_numbers), _ptr(frc), _ptr(target), _ptr(loss),
            _ptr(grad_charge), _ptr(grad_A), _ptr(grad_B), _stream(stream)), "sponge_nb_force_matching_gradient")

    def energy_parts(self):
        import torch
        torch.cuda.synchronize()
        p = self.parts.cpu().numpy()
        return {k: float(v) for k, v in zip(PART_NAMES, p)}

    def shapes(self):
        """The launch shapes the handles run with (a Shapes record)."""
        sh = Shapes()
        _check(load_library().sponge_nb_get_shapes(self.nl.h, self.lj.h, self.pme.h, ctypes.byref(sh)),
               "sponge_nb_get_shapes")
        return sh

    def set_shapes(self, sh):
        """Sets a Shapes record (or a dict of its fields); between steps, not
        inside a graph capture."""
        if isinstance(sh, dict):
            sh = Shapes.from_dict(sh)
        _check(load_library().sponge_nb_set_shapes(self.nl.h, self.lj.h, self.pme.h, ctypes.byref(sh)),
               "sponge_nb_set_shapes")

    def default_shapes(self):
        """The tuned table's shapes for this system (conv_split -1: by plane
        size)."""
        sh = Shapes()
        _check(load_library().sponge_nb_default_shapes(self.n, self.refresh_interval, float(self.sys.cutoff),
                                                        float(self.sys.skin), ctypes.byref(sh)),
               "sponge_nb_default_shapes")
        return sh


class _BatchIo(ctypes.Structure):
    """sponge_batch_io (include/sponge_nb.h)."""
    _fields_ = [(name, _P) for name in (
        "crd", "uint_crd", "charge", "excluded_list_start", "excluded_list", "excluded_numbers",
        "frc", "atom_energy", "atom_virial", "parts")]


class BatchedSystems:
    """Independent systems (FEP windows, replicas) stepped together: every
    stage of the nonbonded step is one launch over all of them
    (sponge_batch_*). Built from NonbondedEngine objects, whose module
    handles and device arrays it uses in place."""

    def __init__(self, engines):
        lib = load_library()
        self.engines = list(engines)
        k = len(self.engines)
        nl = (_P * k)(*[e.nl.h for e in self.engines])
        lj = (_P * k)(*[e.lj.h for e in self.engines])
        pme = (_P * k)(*[e.pme.h for e in self.engines])
        io = (_BatchIo * k)()
        for i, e in enumerate(self.engines):
            io[i] = _BatchIo(_ptr(e.crd), _ptr(e.uint_crd), _ptr(e.charge), _ptr(e.excl_start),
                             _ptr(e.excl_list), _ptr(e.excl_numbers), _ptr(e.frc), _ptr(e.atom_energy),
                             _ptr(e.atom_virial), _ptr(e.parts))
        self._keep = (nl, lj, pme, io)
        h = _P()
        _check(lib.sponge_batch_create(k, ctypes.cast(nl, _P), ctypes.cast(lj, _P), ctypes.cast(pme, _P),
                                       ctypes.cast(io, _P), ctypes.byref(h)), "sponge_batch_create")
        self.h = h

    def neighbor_list_update(self, forced=0, stream=None):
        _check(load_library().sponge_batch_neighbor_list_update(self.h, int(forced), _stream(stream)),
               "sponge_batch_neighbor_list_update")

    def force_step(self, need_energy=0, need_virial=0, from_crd=True, overwrite=True, serial=False, stream=None):
        """serial: every stage on `stream` (else the reciprocal stages run on
        the batch's side stream beside the pair stage)."""
        flags = (OVERWRITE_FORCE if overwrite else 0) | (FROM_CRD if from_crd else 0) | (SERIAL if serial else 0)
        _check(load_library().sponge_batch_force_step(self.h, flags, int(need_energy), int(need_virial),
                                                      _stream(stream)), "sponge_batch_force_step")

    def sync(self):
        _check(load_library().sponge_batch_sync(self.h), "sponge_batch_sync")

    def Clear(self):
        if self.h:
            load_library().sponge_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.Clear()
        except Exception:
            pass
