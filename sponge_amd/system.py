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
TIME_UNIT_PER_PS = 20.455       # SPONGE's internal time unit is 1/20.455 ps (common.cuh)
KB = 0.00198716                 # kcal/mol/K

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


@dataclasses.dataclass
class FepSystem(System):
    """A System with SPONGE's soft-core FEP description (LJ_soft_core.cu:
    1420-1610): LJ types and pair tables of state A (lj_type, lj_A, lj_B) and
    of state B (lj_type_B, lj_A_B, lj_B_B), the subsystem mask, lambda and the
    soft-core parameters. `charge` holds the charges at lambda,
    (1 - lambda) charge_A + lambda charge_B (linear charge scaling)."""
    lj_type_B: np.ndarray = None
    lj_A_B: np.ndarray = None
    lj_B_B: np.ndarray = None
    mask: np.ndarray = None
    charge_A: np.ndarray = None
    charge_B: np.ndarray = None
    lam: float = 0.0
    sc_alpha: float = 0.5
    sc_p: float = 1.0
    sc_sigma: float = 3.0
    sc_sigma_min: float = 0.0

    @property
    def ntypes_B(self) -> int:
        t = len(self.lj_A_B)
        return int(round((math.sqrt(8 * t + 1) - 1) / 2))


def _lb_table(sig, eps):
    """Lorentz-Berthelot pair tables of per-type sigma / epsilon, x12 / x6;
    a type with sigma or epsilon 0 has A = B = 0 with every partner."""
    T = len(sig)
    A = np.zeros(T * (T + 1) // 2, np.float64)
    B = np.zeros_like(A)
    for a in range(T):
        for b in range(a + 1):
            if sig[a] == 0 or sig[b] == 0 or eps[a] == 0 or eps[b] == 0:
                continue
            s = 0.5 * (sig[a] + sig[b])
            e = math.sqrt(eps[a] * eps[b])
            k = a * (a + 1) // 2 + b
            A[k] = 4.0 * e * s ** 12
            B[k] = 4.0 * e * s ** 6
    return (A * 12.0).astype(np.float32), (B * 6.0).astype(np.float32)


WATER_VOLUME = 29.915  # A^3 per water at 1.0 g/cm^3 (18.015 g/mol)


def fep_system(n_water: int, lam: float, seed: int = 11, n_solute: int = 12,
               cutoff: float = 9.0, skin: float = 2.0, name: str = None) -> FepSystem:
    """One solute of `n_solute` atoms decoupled by lambda in a TIP3P box.

    The box holds `n_water` waters at 1.0 g/cm^3 before the solute is placed
    at its centre (vertices of a randomly rotated icosahedron of radius 2.6 A
    for 12 atoms, a random compact cluster otherwise); waters with an atom
    closer than 2.6 A to a solute atom are removed. Solute charges are
    N(0, 0.3) e shifted to zero net charge; LJ sigma U(3.0, 3.8) A, epsilon
    U(0.05, 0.2) kcal/mol, one type per solute atom. State A is the coupled
    solute, state B has zero solute charge and LJ (type `zero`); the solute
    is mask 1, water mask 0; solute-solute pairs are excluded.
    """
    rng = np.random.default_rng(seed)
    L = (n_water * WATER_VOLUME) ** (1.0 / 3.0)
    w = water_box(n_water, L, seed=seed, cutoff=cutoff, skin=skin)
    c = np.full(3, L / 2)
    if n_solute == 12:
        g = (1 + math.sqrt(5)) / 2
        v = np.array([[0, 1, g], [0, -1, g], [0, 1, -g], [0, -1, -g], [1, g, 0], [-1, g, 0],
                      [1, -g, 0], [-1, -g, 0], [g, 0, 1], [-g, 0, 1], [g, 0, -1], [-g, 0, -1]], float)
        v *= 2.6 / np.linalg.norm(v[0])
    else:
        v = rng.normal(size=(n_solute, 3))
        v *= 2.6 / np.linalg.norm(v, axis=1)[:, None]
    sol = c + v @ _random_rotations(rng, 1)[0].T
    # drop overlapping waters (minimum image)
    crd = w.crd.astype(np.float64).reshape(-1, 3, 3)
    d = crd[:, :, None, :] - sol[None, None, :, :]
    d -= L * np.round(d / L)
    keep = ~(np.sqrt((d ** 2).sum(-1)) < 2.6).any(axis=(1, 2))
    keep &= np.linalg.norm(crd[:, 0] - c, axis=1) > 3.0
    nw = int(keep.sum())
    wat = crd[keep].reshape(-1, 3)
    xyz = np.concatenate([wat, sol]).astype(np.float32)
    xyz = np.mod(xyz, np.float32(L)).astype(np.float32)
    xyz[xyz >= np.float32(L)] = 0.0
    N = 3 * nw + n_solute
    # types: O 0, H 1, solute 2..2+n_solute-1, zero type last
    sig = [SPECIES["O"][1], 0.0] + list(rng.uniform(3.0, 3.8, n_solute)) + [0.0]
    eps = [SPECIES["O"][2], 0.0] + list(rng.uniform(0.05, 0.2, n_solute)) + [0.0]
    A, B = _lb_table(sig, eps)
    zero = 2 + n_solute
    tA = np.concatenate([np.tile([0, 1, 1], nw), 2 + np.arange(n_solute)]).astype(np.int32)
    tB = tA.copy()
    tB[3 * nw:] = zero
    qs = rng.normal(0.0, 0.3, n_solute)
    qs -= qs.mean()
    qA = np.concatenate([np.tile([SPECIES["O"][0], SPECIES["H"][0], SPECIES["H"][0]], nw), qs])
    qA = (qA * CHARGE_SCALE).astype(np.float32)
    qB = qA.copy()
    qB[3 * nw:] = 0.0
    q = ((1.0 - lam) * qA.astype(np.float64) + lam * qB.astype(np.float64)).astype(np.float32)
    mask = np.zeros(N, np.int32)
    mask[3 * nw:] = 1
    mass = np.concatenate([np.tile([15.999, 1.008, 1.008], nw), np.full(n_solute, 12.011)]).astype(np.float32)
    excl = [[] for _ in range(N)]
    for k in range(nw):
        o = 3 * k
        excl[o] = [o + 1, o + 2]
        excl[o + 1] = [o + 2]
    for a in range(n_solute):
        excl[3 * nw + a] = [3 * nw + b for b in range(a + 1, n_solute)]
    numbers = np.array([len(e) for e in excl], np.int32)
    start = np.zeros(N, np.int32)
    start[1:] = np.cumsum(numbers)[:-1]
    lst = np.array([j for e in excl for j in e], np.int32)
    return FepSystem(name=name or f"fep{N}_l{lam:.3f}", box=np.full(3, L, np.float32), crd=xyz,
                     charge=q, mass=mass, lj_type=tA, lj_A=A, lj_B=B, excl_start=start,
                     excl_numbers=numbers, excl_list=lst, cutoff=cutoff, skin=skin,
                     waters=nw, lj_type_B=tB, lj_A_B=A.copy(), lj_B_B=B.copy(), mask=mask,
                     charge_A=qA, charge_B=qB, lam=float(lam))


def fep_mixed(fs: FepSystem, n_change: int = 5, n_vanish: int = 5, charge_scale: float = 0.5) -> FepSystem:
    """A FepSystem whose solvent is perturbed too, so that every kind of pair
    of LJ_soft_core.cu:1041-1047 occurs: the oxygens of the first `n_change`
    waters take the first solute type in state B (ordinary pairs whose end
    states differ), those of the next `n_vanish` waters the zero type (soft
    pairs inside one mask), and the charges of all these waters are scaled by
    `charge_scale` in state B (perturbed charges on ordinary pairs). The
    charges at lambda are recomputed."""
    tB = fs.lj_type_B.copy()
    qB = fs.charge_B.copy()
    zero = int(fs.lj_type_B.max())
    for k in range(n_change + n_vanish):
        tB[3 * k] = 2 if k < n_change else zero
        qB[3 * k:3 * k + 3] = (fs.charge_A[3 * k:3 * k + 3].astype(np.float64) * charge_scale).astype(np.float32)
    lam = fs.lam
    q = ((1.0 - lam) * fs.charge_A.astype(np.float64) + lam * qB.astype(np.float64)).astype(np.float32)
    return dataclasses.replace(fs, lj_type_B=tB, charge_B=qB, charge=q, name=fs.name + "_mixed")


def fep_batch(seed: int = 11, window_waters: int = 1000, n_windows: int = 16,
              replica_atoms=(1500, 3000, 4500, 6000, 7500, 9000, 10500, 12000)):
    """The batched-FEP workload: `n_windows` lambda windows (lambda = k/(n-1))
    of one solvated solute, plus replicas of other sizes (about the given atom
    counts) at seeded random lambdas; cutoff 9 A, skin 2 A, PME grids from
    SPONGE's size rule (about 1 A spacing)."""
    rng = np.random.default_rng(seed + 1000)
    out = [fep_system(window_waters, k / (n_windows - 1), seed=seed, name=f"window{k:02d}")
           for k in range(n_windows)]
    for r, n in enumerate(replica_atoms):
        lam = float(rng.uniform(0.0, 1.0))
        out.append(fep_system(max(1, (n - 12) // 3), lam, seed=seed + 1 + r, name=f"replica{r}"))
    return out


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

def water_constraints(sys_: System):
    """Constraint pairs of the rigid waters, in the order SPONGE's CONSTRAIN
    module lists a water's bonds and its H-O-H angle: (O, H1), (O, H2),
    (H1, H2); lengths 0.9572 and 2 * 0.9572 * sin(104.52 / 2) Angstrom."""
    nw = sys_.waters
    o = 3 * np.arange(nw)
    r_hh = 2.0 * TIP3P_ROH * math.sin(math.radians(TIP3P_HOH_DEG) / 2)
    pi = np.stack([o, o, o + 1], 1).reshape(-1).astype(np.int32)
    pj = np.stack([o + 1, o + 2, o + 2], 1).reshape(-1).astype(np.int32)
    pr = np.tile(np.array([TIP3P_ROH, TIP3P_ROH, r_hh], np.float32), nw)
    return pi, pj, pr


def rigid_velocities(sys_: System, temperature: float = 300.0, seed: int = 5) -> np.ndarray:
    """Maxwell-Boltzmann velocities (Angstrom per internal time unit) with
    every water's velocities projected onto rigid-body motion (centre of
    mass velocity + rotation), and the total momentum removed."""
    rng = np.random.default_rng(seed)
    m = sys_.mass.astype(np.float64)
    v = rng.normal(size=(sys_.n, 3)) * np.sqrt(KB * temperature / m)[:, None]
    nw = sys_.waters
    if nw:
        x = sys_.crd[:3 * nw].astype(np.float64).reshape(nw, 3, 3)
        # unwrap each water around its oxygen
        L = sys_.box.astype(np.float64)
        x = x[:, :1] + (x - x[:, :1]) - L * np.round((x - x[:, :1]) / L)
        mw = m[:3 * nw].reshape(nw, 3)
        vw = v[:3 * nw].reshape(nw, 3, 3)
        M = mw.sum(1)
        com = (mw[..., None] * x).sum(1) / M[:, None]
        V = (mw[..., None] * vw).sum(1) / M[:, None]
        r = x - com[:, None]
        Lm = (mw[..., None] * np.cross(r, vw - V[:, None])).sum(1)
        r2 = (r * r).sum(2)
        inertia = (mw * r2).sum(1)[:, None, None] * np.eye(3) - np.einsum("wa,wai,waj->wij", mw, r, r)
        omega = np.linalg.solve(inertia, Lm[..., None])[..., 0]
        vw = V[:, None] + np.cross(omega[:, None], r)
        v[:3 * nw] = vw.reshape(-1, 3)
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v.astype(np.float32)


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

def write_rst7(path: str, title: str, crd, box, vel=None, time: float | None = None) -> None:
    """Writes an AMBER rst7 restart as SPONGE writes one (MD_core.cu:1425-1463):
    the title, `%8d %.3f` atoms and time (only `%8d` when there are no
    velocities), coordinates then velocities as `%12.7f` triples two atoms a
    line, then the box lengths and three 90 degree angles."""
    crd = np.asarray(crd, np.float32).reshape(-1, 3)
    lists = [crd] if vel is None else [crd, np.asarray(vel, np.float32).reshape(-1, 3)]
    with open(path, "w") as f:
        f.write(f"{title}\n")
        f.write(f"{len(crd):8d}" + ("" if vel is None else f" {0.0 if time is None else time:.3f}") + "\n")
        for arr in lists:
            for i in range(0, len(arr), 2):
                f.write("".join(f"{float(x):12.7f}" for x in arr[i:i + 2].reshape(-1)) + "\n")
        b = np.asarray(box, np.float32)
        f.write("".join(f"{float(x):12.7f}" for x in (*b, 90.0, 90.0, 90.0)) + "\n")


def read_rst7(path: str):
    """Reads an AMBER rst7 (fixed 12-column fields, six a line). Returns
    (title, n, time or None, crd, vel or None, box)."""
    with open(path) as f:
        lines = f.read().splitlines()
    title = lines[0]
    head = lines[1].split()
    n = int(head[0])
    time = float(head[1]) if len(head) > 1 else None
    vals = []
    for line in lines[2:]:
        vals.extend(float(line[i:i + 12]) for i in range(0, len(line.rstrip()), 12))
    vals = np.asarray(vals, np.float64)
    crd = vals[:3 * n].reshape(n, 3)
    vel = vals[3 * n:6 * n].reshape(n, 3) if time is not None else None
    rest = vals[6 * n:] if time is not None else vals[3 * n:]
    return title, n, time, crd, vel, rest[:3]


def write_sponge_inputs(sys_: System, directory: str, step_limit: int = 1,
                        write_mdout_interval: int = 1, refresh_interval: int = 0,
                        extra: dict | None = None, velocities: np.ndarray | None = None,
                        dt_ps: float | None = None, rst7: bool = False) -> str:
    """Writes `sys_` as SPONGE text inputs plus an mdin and returns the mdin path.

    Formats are the ones MD_core.cu reads: coordinate_in_file (`N start_time`,
    N lines `x y z`, then the box; :1212-1256), charge_in_file (`N`, then N
    charges already x18.2223; :1053-1080), exclude_in_file (`N total`, then
    per atom `count j...`; :308-345) and LJ_in_file (`N T`, A and B of the
    T(T+1)/2 pairs before the x12 / x6 scaling, then N types;
    Lennard_Jones_force.cu:1076-1108).

    With `velocities` the molecular-dynamics inputs follow: velocity_in_file
    (`N`, then N lines; MD_core.cu:864-890), mass_in_file (`N`, then N masses;
    :922-949), constrain_in_file (the rigid waters' constraint pairs: a count,
    then `i j r`; constrain/constrain.cu:27-64), mode NVE and dt (ps).

    With `rst7` the coordinates (and the velocities, when given) go into one
    AMBER rst7 file named by amber_rst7 (MD_core.cu:1258-1347) instead.
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
    md_keys = {}
    if velocities is not None:
        v = np.asarray(velocities, np.float32).reshape(-1, 3)
        with open(p("velocity.txt"), "w") as f:
            f.write(f"{sys_.n}\n")
            for x, y, z in v:
                f.write(f"{x:.9g} {y:.9g} {z:.9g}\n")
        with open(p("mass.txt"), "w") as f:
            f.write(f"{sys_.n}\n")
            f.write("\n".join(f"{m:.9g}" for m in sys_.mass.astype(np.float32)) + "\n")
        pi, pj, pr = water_constraints(sys_)
        with open(p("constrain.txt"), "w") as f:
            f.write(f"{len(pi)}\n")
            for a, b, r in zip(pi, pj, pr):
                f.write(f"{int(a)} {int(b)} {float(r):.9g}\n")
        md_keys = {"velocity_in_file": p("velocity.txt"), "mass_in_file": p("mass.txt"),
                   "constrain_in_file": p("constrain.txt"), "mode": "NVE",
                   "dt": 0.001 if dt_ps is None else dt_ps}
    crd_keys = {"coordinate_in_file": p("coordinate.txt")}
    if rst7:
        write_rst7(p("system.rst7"), sys_.name, sys_.crd, sys_.box, velocities,
                   None if velocities is None else 0.0)
        crd_keys = {"amber_rst7": p("system.rst7")}
        md_keys.pop("velocity_in_file", None)
    keys = {
        **crd_keys,
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
    keys.update(md_keys)
    keys.update(extra or {})
    mdin = p("mdin.txt")
    with open(mdin, "w") as f:
        f.write(f"{sys_.name}\n")
        for k, v in keys.items():
            f.write(f"{k} = {v}\n")
    return mdin


def write_fe_inputs(fs: FepSystem, directory: str, frames, boxes=None, charge_perturbed: int = 1,
                    extra: dict | None = None) -> str:
    """Writes a FepSystem and a trajectory as the inputs of SPONGE_FEP /
    SPONGE_TI (and of sponge_fe_driver) and returns the mdin path.

    Files: LJ_soft_core_in_file (`N TA TB`, AA and AB of the A pairs and BA
    and BB of the B pairs before the x12 / x6 scaling, then per atom
    `typeA typeB`; LJ_soft_core.cu:1483-1518), subsys_division_in_file
    (`N`, then N masks), charge_in_file (the charges at lambda),
    chargeA_in_file / chargeB_in_file (TI_core.cu:431-463), exclude_in_file,
    crd (float32 xyz frames, as mdcrd.dat) and box (`Lx Ly Lz 90 90 90` per
    frame); `frames` is a sequence of (N, 3) arrays and `boxes` of 3-vectors
    (default: the system's box for every frame)."""
    import os
    os.makedirs(directory, exist_ok=True)
    p = lambda name: os.path.join(directory, name)
    frames = [np.asarray(f, np.float32).reshape(-1, 3) for f in frames]
    boxes = [fs.box] * len(frames) if boxes is None else boxes
    with open(p("LJ_soft_core.txt"), "w") as f:
        f.write(f"{fs.n} {fs.ntypes} {fs.ntypes_B}\n")
        f.write(" ".join(f"{float(a) / 12.0:.17g}" for a in fs.lj_A) + "\n")
        f.write(" ".join(f"{float(b) / 6.0:.17g}" for b in fs.lj_B) + "\n")
        f.write(" ".join(f"{float(a) / 12.0:.17g}" for a in fs.lj_A_B) + "\n")
        f.write(" ".join(f"{float(b) / 6.0:.17g}" for b in fs.lj_B_B) + "\n")
        f.write("\n".join(f"{int(a)} {int(b)}" for a, b in zip(fs.lj_type, fs.lj_type_B)) + "\n")
    for name, arr, fmt in (("subsys_division.txt", fs.mask, "{:d}"), ("charge.txt", fs.charge, "{:.9g}"),
                           ("chargeA.txt", fs.charge_A, "{:.9g}"), ("chargeB.txt", fs.charge_B, "{:.9g}")):
        with open(p(name), "w") as f:
            f.write(f"{fs.n}\n")
            f.write("\n".join(fmt.format(x.item()) for x in arr) + "\n")
    with open(p("exclude.txt"), "w") as f:
        f.write(f"{fs.n} {len(fs.excl_list)}\n")
        for i in range(fs.n):
            s, k = int(fs.excl_start[i]), int(fs.excl_numbers[i])
            f.write(" ".join([str(k)] + [str(int(j)) for j in fs.excl_list[s:s + k]]) + "\n")
    with open(p("mdcrd.dat"), "wb") as f:
        for fr in frames:
            f.write(fr.astype(np.float32).tobytes())
    with open(p("box.txt"), "w") as f:
        for b in boxes:
            f.write(f"{float(b[0]):.9g} {float(b[1]):.9g} {float(b[2]):.9g} 90.0 90.0 90.0\n")
    keys = {
        "atom_numbers": fs.n,
        "frame_numbers": len(frames),
        "crd": p("mdcrd.dat"),
        "box": p("box.txt"),
        "charge_pertubated": charge_perturbed,
        "charge_in_file": p("charge.txt"),
        "chargeA_in_file": p("chargeA.txt"),
        "chargeB_in_file": p("chargeB.txt"),
        "subsys_division_in_file": p("subsys_division.txt"),
        "exclude_in_file": p("exclude.txt"),
        "LJ_soft_core_in_file": p("LJ_soft_core.txt"),
        "lambda_lj": fs.lam,
        "soft_core_alpha": fs.sc_alpha,
        "soft_core_powfer": fs.sc_p,
        "soft_core_sigma": fs.sc_sigma,
        "soft_core_sigma_min": fs.sc_sigma_min,
        "cutoff": fs.cutoff,
        "skin": fs.skin,
        "PME_Direct_Tolerance": fs.pme_tolerance,
        "float32ene": p("float32ene.dat"),
        "TI": p("TI.txt"),
    }
    for axis, k in zip("xyz", fs.fft):
        if k > 0:
            keys[f"PME_fft{axis}"] = k
    keys.update(extra or {})
    mdin = p("mdin.txt")
    with open(mdin, "w") as f:
        f.write(f"{fs.name}\n")
        for k, v in keys.items():
            f.write(f"{k} = {v}\n")
    return mdin
