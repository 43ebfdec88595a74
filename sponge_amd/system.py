"""Seeded synthetic periodic water(+ion) boxes in SPONGE's internal units.

Every test, the benchmark and the CPU baseline build their inputs here, so
CPU and GPU always see the same system for a given (config, seed).

Units follow SPONGE: lengths in Angstrom, energies in kcal/mol, charges stored
pre-multiplied by 18.2223 (MD_core/MD_core.cu:1051-1080 reads them that way),
LJ A/B stored pre-multiplied by 12 and 6 in a lower-triangular pair table
(Lennard_Jones_force/Lennard_Jones_force.cu:1092,1097; pair index at :106-112).
Exclusions are an upper-triangular CSR sorted per atom
(MD_core/MD_core.cu:284-483, sort at :441).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

CHARGE_SCALE = 18.2223          # sqrt(kcal*A/mol) per elementary charge
TIP3P_ROH = 0.9572
TIP3P_HOH_DEG = 104.52

# (name, charge e, sigma A, epsilon kcal/mol)
SPECIES = {
    "O": (-0.834, 3.15061, 0.1521),
    "H": (0.417, 0.0, 0.0),
    "Na": (1.0, 2.43, 0.0874),
    "Cl": (-1.0, 4.40, 0.1),
}


@dataclasses.dataclass
class System:
    """One periodic system; arrays are host numpy in original atom order."""
    name: str
    box: np.ndarray               # (3,) float32, Angstrom
    crd: np.ndarray               # (N,3) float32, wrapped into [0,L)
    charge: np.ndarray            # (N,) float32, x18.2223
    mass: np.ndarray              # (N,) float32
    lj_type: np.ndarray           # (N,) int32
    lj_A: np.ndarray              # (T(T+1)/2,) float32, already x12
    lj_B: np.ndarray              # (T(T+1)/2,) float32, already x6
    excl_start: np.ndarray        # (N,) int32
    excl_numbers: np.ndarray      # (N,) int32
    excl_list: np.ndarray         # (E,) int32
    cutoff: float = 10.0
    skin: float = 2.0
    pme_tolerance: float = 1e-5
    fft: tuple = (-1, -1, -1)     # -1 = SPONGE's automatic size
    waters: int = 0               # first 3*waters atoms are O,H1,H2 triples

    @property
    def n(self) -> int:
        return int(self.crd.shape[0])

    @property
    def ntypes(self) -> int:
        t = len(self.lj_A)
        return int(round((math.sqrt(8 * t + 1) - 1) / 2))

    def pair_table_square(self):
        """Full T x T (A, B) tables expanded from the triangular storage."""
        T = self.ntypes
        A = np.zeros((T, T), np.float32)
        B = np.zeros((T, T), np.float32)
        for a in range(T):
            for b in range(T):
                hi, lo = max(a, b), min(a, b)
                k = hi * (hi + 1) // 2 + lo
                A[a, b] = self.lj_A[k]
                B[a, b] = self.lj_B[k]
        return A, B


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    R = np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)
    return R.reshape(n, 3, 3)


def lj_tables(species_order):
    """Lorentz-Berthelot A=4 eps sigma^12, B=4 eps sigma^6, stored x12 / x6."""
    T = len(species_order)
    A = np.zeros(T * (T + 1) // 2, np.float64)
    B = np.zeros_like(A)
    for a in range(T):
        for b in range(a + 1):
            _, sa, ea = SPECIES[species_order[a]]
            _, sb, eb = SPECIES[species_order[b]]
            k = a * (a + 1) // 2 + b
            if sa == 0.0 or sb == 0.0 or ea == 0.0 or eb == 0.0:
                continue
            s = 0.5 * (sa + sb)
            e = math.sqrt(ea * eb)
            A[k] = 4.0 * e * s ** 12
            B[k] = 4.0 * e * s ** 6
    return (A * 12.0).astype(np.float32), (B * 6.0).astype(np.float32)


def water_box(n_water: int, box: float, n_na: int = 0, n_cl: int = 0,
              seed: int = 2024, cutoff: float = 10.0, skin: float = 2.0,
              fft=(-1, -1, -1), name: str = "water") -> System:
    """Rigid TIP3P waters (+ ions) on a jittered cubic lattice.

    Oxygen (and ion) sites sit on an n^3 lattice (n^3 >= waters+ions) of
    spacing box/n, jittered uniformly by +-0.15 spacing; a seeded random
    subset of sites is used, the last n_na+n_cl of them become ions.
    Each water gets a uniformly random orientation.
    """
    rng = np.random.default_rng(seed)
    nsite = n_water + n_na + n_cl
    n = max(1, int(round(nsite ** (1.0 / 3.0))))
    while n ** 3 < nsite:
        n += 1
    a = box / n
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                             indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pick = rng.permutation(len(g))[:nsite]
    sites = (g[pick] + 0.5) * a + rng.uniform(-0.15, 0.15, (nsite, 3)) * a
    O = sites[:n_water]
    ions = sites[n_water:]
    R = _random_rotations(np.random.default_rng(seed + 1), n_water)
    th = math.radians(TIP3P_HOH_DEG) / 2
    h1 = np.array([TIP3P_ROH * math.sin(th), 0.0, TIP3P_ROH * math.cos(th)])
    h2 = np.array([-TIP3P_ROH * math.sin(th), 0.0, TIP3P_ROH * math.cos(th)])
    H1 = O + R @ h1
    H2 = O + R @ h2
    crd = np.concatenate([np.stack([O, H1, H2], 1).reshape(-1, 3), ions])
    crd = np.mod(crd, box)
    crd = crd.astype(np.float32)
    crd[crd >= np.float32(box)] = 0.0

    order = ["O", "H", "Na", "Cl"]
    types = np.concatenate([np.tile([0, 1, 1], n_water),
                            np.full(n_na, 2), np.full(n_cl, 3)]).astype(np.int32)
    q_e = np.array([SPECIES[order[t]][0] for t in types])
    charge = (q_e * CHARGE_SCALE).astype(np.float32)
    mass_tab = {0: 15.999, 1: 1.008, 2: 22.990, 3: 35.45}
    mass = np.array([mass_tab[t] for t in types], np.float32)
    A, B = lj_tables(order)

    N = len(types)
    excl_numbers = np.zeros(N, np.int32)
    excl_numbers[0:3 * n_water:3] = 2
    excl_numbers[1:3 * n_water:3] = 1
    excl_start = np.zeros(N, np.int32)
    excl_start[1:] = np.cumsum(excl_numbers)[:-1]
    lst = []
    for w in range(n_water):
        o = 3 * w
        lst += [o + 1, o + 2, o + 2]
    excl_list = np.array(lst, np.int32)
    return System(name=name, box=np.full(3, box, np.float32), crd=crd,
                  charge=charge, mass=mass, lj_type=types, lj_A=A, lj_B=B,
                  excl_start=excl_start, excl_numbers=excl_numbers,
                  excl_list=excl_list, cutoff=cutoff, skin=skin,
                  fft=tuple(fft), waters=n_water)


def config(idx: int, seed: int = 2024) -> System:
    """Benchmark workloads by index (water boxes at 1.0 g/cm^3).

    0: 216 waters, 18.64 A box, rc 8 / skin 2, PME grid auto (20^3)
    1: 7,850 waters + 10 Na + 10 Cl, 61.9 A box, rc 10 / skin 2, 64^3
    2: 32,000 waters + 60 Na + 60 Cl, 98.6 A box, rc 10 / skin 2, 108^3
    """
    if idx == 0:
        return water_box(216, 18.64, seed=seed, cutoff=8.0, skin=2.0,
                         name="water648")
    if idx == 1:
        return water_box(7850, 61.9, 10, 10, seed=seed, cutoff=10.0, skin=2.0,
                         name="water23570")
    if idx == 2:
        return water_box(32000, 98.6, 60, 60, seed=seed, cutoff=10.0,
                         skin=2.0, fft=(108, 108, 108), name="water96120")
    raise ValueError(f"unknown config {idx}")


def small_random(n: int, box, seed: int = 7, cutoff: float = 4.0,
                 skin: float = 1.0, n_types: int = 3, excl_frac: float = 0.2,
                 fft=(-1, -1, -1)) -> System:
    """Random point charges with random LJ types and random exclusions.

    Used for edge-case tests (small / non-cubic boxes, uneven types)."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, np.float32).reshape(3)
    crd = (rng.uniform(0, 1, (n, 3)) * box).astype(np.float32)
    crd = np.where(crd >= box, 0, crd).astype(np.float32)
    q = rng.normal(0, 0.5, n)
    q -= q.mean()
    charge = (q * CHARGE_SCALE).astype(np.float32)
    types = rng.integers(0, n_types, n).astype(np.int32)
    T = n_types
    A = np.zeros(T * (T + 1) // 2, np.float32)
    B = np.zeros_like(A)
    sig = rng.uniform(2.0, 3.5, T)
    eps = rng.uniform(0.05, 0.2, T)
    for a in range(T):
        for b in range(a + 1):
            s = 0.5 * (sig[a] + sig[b])
            e = math.sqrt(eps[a] * eps[b])
            k = a * (a + 1) // 2 + b
            A[k] = 12.0 * 4 * e * s ** 12
            B[k] = 6.0 * 4 * e * s ** 6
    excl = [[] for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, min(n, i + 4)):
            if rng.uniform() < excl_frac:
                excl[i].append(j)
    numbers = np.array([len(e) for e in excl], np.int32)
    start = np.zeros(n, np.int32)
    start[1:] = np.cumsum(numbers)[:-1]
    lst = np.array([j for e in excl for j in e], np.int32)
    return System(name=f"random{n}", box=box, crd=crd, charge=charge,
                  mass=np.ones(n, np.float32), lj_type=types, lj_A=A, lj_B=B,
                  excl_start=start, excl_numbers=numbers, excl_list=lst,
                  cutoff=cutoff, skin=skin, fft=tuple(fft))


# ---------------------------------------------------------------------------
# SPONGE host-side conventions restated (pure arithmetic, used by host code)
# ---------------------------------------------------------------------------

def check_2357(n: int) -> bool:
    """common.cu:467-511 Check_2357_Factor."""
    if n <= 0:
        return False
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def fft_size(length: float) -> int:
    """common.cu:513-534 Get_Fft_Patameter."""
    t = (int(math.ceil(np.float32(length) + 3)) >> 2) << 2
    for lo, hi, snap in ((60, 68, 64), (120, 136, 128), (240, 272, 256),
                         (480, 544, 512), (960, 1088, 1024)):
        if lo <= t <= hi:
            t = snap
            break
    while not check_2357(t):
        t += 4
    return t


# ---------------------------------------------------------------------------
# SPONGE input files
# ---------------------------------------------------------------------------

def write_sponge_inputs(sys_: System, directory: str, step_limit: int = 1,
                        write_mdout_interval: int = 1, refresh_interval: int = 0,
                        extra: dict | None = None) -> str:
    """Writes `sys_` as SPONGE text inputs plus an mdin and returns the mdin path.

    Formats are the ones MD_core.cu reads: coordinate_in_file (`N start_time`,
    N lines `x y z`, then the box; :1212-1256), charge_in_file (`N`, then N
    charges already x18.2223; :1053-1080), exclude_in_file (`N total`, then
    per atom `count j...`; :308-345) and LJ_in_file (`N T`, A and B of the
    T(T+1)/2 pairs before the x12 / x6 scaling, then N types;
    Lennard_Jones_force.cu:1076-1108).
    """
    import os
    os.makedirs(directory, exist_ok=True)
    p = lambda name: os.path.join(directory, name)
    with open(p("coordinate.txt"), "w") as f:
        f.write(f"{sys_.n} 0.0\n")
        for x, y, z in sys_.crd.astype(np.float32):
            f.write(f"{x:.9g} {y:.9g} {z:.9g}\n")
        f.write(f"{sys_.box[0]:.9g} {sys_.box[1]:.9g} {sys_.box[2]:.9g} 90.0 90.0 90.0\n")
    with open(p("charge.txt"), "w") as f:
        f.write(f"{sys_.n}\n")
        f.write("\n".join(f"{q:.9g}" for q in sys_.charge.astype(np.float32)) + "\n")
    with open(p("exclude.txt"), "w") as f:
        f.write(f"{sys_.n} {len(sys_.excl_list)}\n")
        for i in range(sys_.n):
            s, k = int(sys_.excl_start[i]), int(sys_.excl_numbers[i])
            f.write(" ".join([str(k)] + [str(int(j)) for j in sys_.excl_list[s:s + k]]) + "\n")
    with open(p("LJ.txt"), "w") as f:
        f.write(f"{sys_.n} {sys_.ntypes}\n")
        f.write(" ".join(f"{float(a) / 12.0:.17g}" for a in sys_.lj_A) + "\n")
        f.write(" ".join(f"{float(b) / 6.0:.17g}" for b in sys_.lj_B) + "\n")
        f.write("\n".join(str(int(t)) for t in sys_.lj_type) + "\n")
    keys = {
        "coordinate_in_file": p("coordinate.txt"),
        "charge_in_file": p("charge.txt"),
        "exclude_in_file": p("exclude.txt"),
        "LJ_in_file": p("LJ.txt"),
        "cutoff": sys_.cutoff,
        "skin": sys_.skin,
        "step_limit": step_limit,
        "write_mdout_interval": write_mdout_interval,
        "neighbor_list_refresh_interval": refresh_interval,
        "PME_Direct_Tolerance": sys_.pme_tolerance,
    }
    for axis, k in zip("xyz", sys_.fft):
        if k > 0:
            keys[f"PME_fft{axis}"] = k
    keys.update(extra or {})
    mdin = p("mdin.txt")
    with open(mdin, "w") as f:
        f.write(f"{sys_.name}\n")
        for k, v in keys.items():
            f.write(f"{k} = {v}\n")
    return mdin
