"""The five synthetic workloads of the line-by-line path and their generator.

Line lists come from the library's seeded counter-based generator
(hk_synth_lines), so every host regenerates identical lists.  Per line:
nu U[0, 30000) cm^-1, log10 S U[-30, -19] (S in cm/molecule, stored as the
line-file strength with Q = 1), E'' U[0, 20000) cm^-1, gamma_air = gamma_self
U[0.01, 0.1] cm^-1/atm, n U[0.3, 0.8], A = 0, delta = 0.  The heavy-tailed
kind draws gamma log-uniform in [1e-4, 1] and gives 5 % of lines
log10 S U[-20, -17].  The partition function is constant (Q = 1 at every T).
Pressures are given in bar and converted with 1 bar = 0.986923 atm
(define.h:37).
"""
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

NA = 6.0221412927e23
POBAR = 0.986923
SEED = 0x48454C494F534B00


@dataclass
class Workload:
    name: str
    lines: List[int]                 # lines per species
    masses: List[float]              # amu per species
    kind: int                        # generator kind (0 standard, 1 heavy-tailed)
    tp: List[Tuple[float, float]]    # (T [K], P [bar])
    numin: float
    numax: float
    dnu: float
    cut: float                       # cm^-1 (cut mode 0)
    nbins: int
    nC: int
    seed: int = SEED
    note: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def nx(self):
        return int((self.numax - self.numin) / self.dnu + 0.5)


def _c2_tp():
    k = np.arange(64)
    T = 100.0 * 30.0 ** (k / 63.0)
    P = 1e-6 * 10.0 ** (9.0 * k / 63.0)
    return [(float(a), float(b)) for a, b in zip(T, P)]


def _c4_tp():
    k = np.arange(16)
    T = 300.0 * 10.0 ** (k / 15.0)
    P = 1e-2 * 10.0 ** (4.0 * k / 15.0)
    return [(float(a), float(b)) for a, b in zip(T, P)]


WORKLOADS = {
    "c0": Workload("plain-fp64-small", [100_000], [18.0], 0, [(1500.0, 1.0 / POBAR)],
                   0.0, 30000.0, 0.01, 100.0, 300, 20, SEED + 0,
                   note="1e5 lines, 3e6 points, T=1500 K, P=1 atm, cut 100"),
    "c1": Workload("precision-at-scale", [10_000_000], [18.0], 0, [(2000.0, 10.0)],
                   0.0, 30000.0, 0.001, 25.0, 300, 20, SEED + 1,
                   note="1e7 lines, 3e7 points, T=2000 K, P=10 bar, cut 25"),
    "c2": Workload("batched-uneven", [10_000, 100_000, 500_000, 1_000_000, 3_000_000, 5_000_000],
                   [2.0, 16.0, 18.0, 28.0, 44.0, 48.0], 0, _c2_tp(),
                   0.0, 30000.0, 0.01, 100.0, 300, 20, SEED + 2,
                   note="6 species x 64 (T,P), 3e6 points, cut 100"),
    "c3": Workload("differentiation", [1_000_000], [18.0], 0, [(1200.0, 0.1)],
                   0.0, 10000.0, 0.01, 50.0, 100, 20, SEED + 3,
                   note="1e6 lines, 1e6 points, T=1200 K, P=0.1 bar, cut 50"),
    "c4": Workload("mixed-width-stress", [50_000_000], [18.0], 1, _c4_tp(),
                   0.0, 40000.0, 0.01, 25.0, 400, 20, SEED + 4,
                   note="5e7 heavy-tailed lines x 16 (T,P), 4e6 points, cut 25"),
}


def species_arrays(w: Workload, s: int, n: int = None):
    """Unsorted numpy line list of species s of workload w (first n lines)."""
    from .engine import synth_lines
    n = w.lines[s] if n is None else n
    d = synth_lines(w.seed, s, n, w.kind)
    return dict(nu=d["nu"], S=d["S"], EL=d["EL"], vy=d["gamma"], n=d["n"],
                iso_mass=(w.masses[s] / NA,), iso_ab=(1.0,), iso_sscale=(1.0,))


def oracle_lines(arr):
    """The same list in the form the FP64 oracle takes, sorted by nu."""
    o = np.argsort(arr["nu"], kind="stable")
    n = o.size
    return dict(nu=arr["nu"][o], S=arr["S"][o], EL=arr["EL"][o], A=np.zeros(n),
                delta=np.zeros(n), vy=arr["vy"][o], n=arr["n"][o], iso=None,
                iso_mass=np.array(arr["iso_mass"]), iso_ab=np.array(arr["iso_ab"]),
                iso_sscale=np.array(arr["iso_sscale"]), iso_Q=np.array([1.0]))
