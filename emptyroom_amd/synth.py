"""Seeded synthetic frameworks for the five benchmark workloads.

The random stream is SplitMix64 (Steele, Lea, Flood, OOPSLA'14) written out
here with numpy integers, so every framework is a pure function of its seed on
any machine.  Cell geometry follows src/Framework.cc:113-142 and mixing
src/load_fast_particle_f_array.h:35-36.  The replication factors of the
synthetic workloads come from their own rule: the smallest integer r_i with
r_i * w_i > 2 r_c, w_i the perpendicular width of the cell (`replication`;
the image half-range must reach r_c past every face of the cell).  This is
not the doubling rule of julia_codes/framework.jl:162-211, which the drop-ins
use through er_replication_factors; the two can differ by one image layer.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

MASK = (1 << 64) - 1
MASTER_SEED = 20261202

# guests (eps K, sigma A) of the workloads
GUESTS = {"CH4": (148.0, 3.73), "Xe": (221.0, 4.10), "Kr": (166.4, 3.636)}


def splitmix64(state: int) -> Tuple[int, int]:
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


class Rng:
    """SplitMix64 stream; uniform() gives (z >> 11) * 2^-53 in [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            s = np.uint64(self.state) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = s
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & MASK
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def scalar(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return float(self.uniform(1, lo, hi)[0])


def framework_seed(config: int, index: int, master: int = MASTER_SEED) -> int:
    _, a = splitmix64((master * 1000003 + config) & MASK)
    _, b = splitmix64((a ^ (index * 0x632BE59BD9B4E019)) & MASK)
    return b


def cell_matrix(a, b, c, alpha, beta, gamma) -> Tuple[np.ndarray, float]:
    """(t_matrix, volume) as src/Framework.cc:113-132 (angles in degrees)."""
    al, be, ga = (x * math.pi / 180.0 for x in (alpha, beta, gamma))
    v = math.sqrt(1.0 - math.cos(al) ** 2 - math.cos(be) ** 2 - math.cos(ga) ** 2
                  + 2 * math.cos(al) * math.cos(be) * math.cos(ga))
    t = np.array([[a, b * math.cos(ga), c * math.cos(be)],
                  [0.0, b * math.sin(ga), c * (math.cos(al) - math.cos(be) * math.cos(ga))
                   / math.sin(ga)],
                  [0.0, 0.0, c * v / math.sin(ga)]])
    return t, v * a * b * c


def perp_widths(t: np.ndarray) -> np.ndarray:
    inv = np.linalg.inv(t)
    return 1.0 / np.linalg.norm(inv, axis=1)


def replication(t: np.ndarray, rc: float) -> List[int]:
    w = perp_widths(t)
    return [int(math.floor(2.0 * rc / wi)) + 1 for wi in w]


@dataclass
class SynthFramework:
    name: str
    cell: Tuple[float, float, float, float, float, float]
    t: np.ndarray
    volume: float
    frac: np.ndarray            # (N, 3)
    types: np.ndarray           # (N,) int
    type_eps: np.ndarray        # (ntypes,)
    type_sig: np.ndarray        # (ntypes,)
    guest: str
    rep: List[int]
    rc: float = 12.5
    spacing: float = 0.5
    n_insertions: int = 0
    temperature: float = 298.0
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def n_atoms(self) -> int:
        return len(self.frac)

    @property
    def rc2(self) -> float:
        return self.rc * self.rc

    def mixed(self) -> Tuple[np.ndarray, np.ndarray]:
        """Per-atom (eps_ij, sig2_ij) by Lorentz-Berthelot."""
        eg, sg = GUESTS[self.guest]
        eps = np.sqrt(eg * self.type_eps[self.types])
        sig2 = ((sg + self.type_sig[self.types]) / 2.0) ** 2
        return eps, sig2

    def particles(self) -> np.ndarray:
        """(N, 5) FrameworkParticle rows."""
        eps, sig2 = self.mixed()
        return np.column_stack([self.frac, eps, sig2])

    def grid_n(self) -> List[int]:
        """N_i = ceil(L_i / spacing) as src/writegrid.cu:110-112."""
        return [int(math.ceil(L / self.spacing)) for L in self.cell[:3]]

    def density(self, masses: dict) -> float:
        return 0.0


def _random_cell(rng: Rng, lmin, lmax, amin, amax, vmin=0.2):
    while True:
        a, b, c = rng.uniform(3, lmin, lmax)
        al, be, ga = rng.uniform(3, amin, amax)
        cal, cbe, cga = (math.cos(x * math.pi / 180) for x in (al, be, ga))
        vf = 1 - cal * cal - cbe * cbe - cga * cga + 2 * cal * cbe * cga
        if vf > vmin * vmin:
            return float(a), float(b), float(c), float(al), float(be), float(ga)


def _types(rng: Rng, ntypes: int):
    eps = rng.uniform(ntypes, 20.0, 250.0)
    sig = rng.uniform(ntypes, 2.5, 4.0)
    return eps, sig


def _min_sep_positions(rng: Rng, t: np.ndarray, n: int, dmin: float) -> np.ndarray:
    pts = np.empty((0, 3))
    while len(pts) < n:
        cand = rng.uniform(3).reshape(1, 3)
        if len(pts):
            d = pts - cand
            d -= np.round(d)
            r = np.linalg.norm(d @ t.T, axis=1)
            if r.min() < dmin:
                continue
        pts = np.vstack([pts, cand])
    return pts


def config1(seed: int = None) -> SynthFramework:
    """25 A cubic cell, 400 atoms (min separation 1.5 A), 4 types, CH4, r=(1,1,1)."""
    seed = framework_seed(1, 0) if seed is None else seed
    rng = Rng(seed)
    t, v = cell_matrix(25.0, 25.0, 25.0, 90.0, 90.0, 90.0)
    te, ts = _types(rng, 4)
    frac = _min_sep_positions(rng, t, 400, 1.5)
    types = (rng.u64(400) % np.uint64(4)).astype(np.int64)
    return SynthFramework("cfg1", (25.0, 25.0, 25.0, 90.0, 90.0, 90.0), t, v, frac, types, te,
                          ts, "CH4", [1, 1, 1], spacing=0.5, n_insertions=100000, seed=seed)


def config2(seed: int = None, n_atoms: int = 3000) -> SynthFramework:
    """Triclinic 12-30 A / 70-110 deg, 3000 atoms, 6 types, CH4, 0.1 A grid."""
    seed = framework_seed(2, 0) if seed is None else seed
    rng = Rng(seed)
    cell = _random_cell(rng, 12.0, 30.0, 70.0, 110.0)
    t, v = cell_matrix(*cell)
    te, ts = _types(rng, 6)
    frac = rng.uniform(3 * n_atoms).reshape(n_atoms, 3)
    types = (rng.u64(n_atoms) % np.uint64(6)).astype(np.int64)
    return SynthFramework("cfg2", cell, t, v, frac, types, te, ts, "CH4", replication(t, 12.5),
                          spacing=0.1, n_insertions=10_000_000, seed=seed)


def config3_framework(index: int, seed: int = None, config: int = 3) -> SynthFramework:
    """One framework of the batching mix: N log-uniform [50,5000], L U[8,40],
    angles U[60,120], spacing in {0.1,0.2,0.5}, guest in {CH4,Xe,Kr}, 1e6 insertions."""
    seed = framework_seed(config, index) if seed is None else seed
    rng = Rng(seed)
    n = int(round(math.exp(rng.scalar(math.log(50.0), math.log(5000.0)))))
    cell = _random_cell(rng, 8.0, 40.0, 60.0, 120.0)
    t, v = cell_matrix(*cell)
    te, ts = _types(rng, 6)
    spacing = [0.1, 0.2, 0.5][int(rng.u64(1)[0] % np.uint64(3))]
    guest = ["CH4", "Xe", "Kr"][int(rng.u64(1)[0] % np.uint64(3))]
    frac = rng.uniform(3 * n).reshape(n, 3)
    types = (rng.u64(n) % np.uint64(6)).astype(np.int64)
    return SynthFramework(f"cfg{config}_{index:03d}", cell, t, v, frac, types, te, ts, guest,
                          replication(t, 12.5), spacing=spacing, n_insertions=1_000_000,
                          seed=seed)


def config3(count: int = 256) -> List[SynthFramework]:
    return [config3_framework(i) for i in range(count)]


def config4_stream(rate: float, n_requests: int, seed: int = None):
    """Poisson arrivals (seconds) of mixed grid / Henry requests, frameworks as in
    the batching mix, T uniform in [250, 350] K."""
    seed = framework_seed(4, 0) if seed is None else seed
    rng = Rng(seed)
    gaps = -np.log1p(-rng.uniform(n_requests)) / rate
    arrivals = np.cumsum(gaps)
    kinds = (rng.u64(n_requests) % np.uint64(2)).astype(np.int64)  # 0 grid, 1 Henry
    temps = rng.uniform(n_requests, 250.0, 350.0)
    return arrivals, kinds, temps


def config5_framework(index: int) -> SynthFramework:
    """Framework `index` of the pocket-blocking workload: drawn from the batching
    distribution, on a 0.2 A grid with 1e7 trilinear insertions."""
    f = config3_framework(index, config=5)
    f.spacing = 0.2
    f.n_insertions = 10_000_000
    return f


def config5(count: int = 64) -> List[SynthFramework]:
    """Frameworks from the batching distribution on a 0.2 A grid with 1e7 trilinear insertions."""
    return [config5_framework(i) for i in range(count)]


def small_framework(seed: int, n_atoms: int = 30, cell=(11.0, 12.0, 13.0, 80.0, 95.0, 105.0),
                    guest: str = "CH4", ntypes: int = 3, rc: float = 12.5, rep=None,
                    spacing: float = 1.0) -> SynthFramework:
    """A small triclinic framework for parity tests."""
    rng = Rng(seed)
    t, v = cell_matrix(*cell)
    te, ts = _types(rng, ntypes)
    frac = rng.uniform(3 * n_atoms).reshape(n_atoms, 3)
    types = (rng.u64(n_atoms) % np.uint64(ntypes)).astype(np.int64)
    rep = replication(t, rc) if rep is None else list(rep)
    return SynthFramework(f"small_{seed}", tuple(cell), t, v, frac, types, te, ts, guest, rep,
                          rc=rc, spacing=spacing, seed=seed)
