"""Workload definitions of the PIC time step (physics shared by all sizes).

Physics (plain electrostatic PIC, no collisions, absorbing walls):
  cell size dx = dy = 1e-4 m; walls phi = 0; electrons and singly charged
  ions of 40 amu; Maxwellian velocities, Te = 2 eV, Ti = 0.03 eV; density
  n0 = 1e16 m^-3 (Debye length 1.05e-4 m, about one cell); dt from
  omega_pe * dt = 0.1 (dt = 1.77e-11 s).

The reference's weighting factor W = actual * area / (sim * dx^2 * dy^2)
makes final_mesh a number density, so each species' `actual` is n0.

Workloads:
  plain_cpu_128   128 x 128 cells, 100k + 100k, Bz = 0.01 T, seed 1, FP64.
  electrode_512   512 x 512 cells, centred 64 x 64-cell electrode at +100 V,
                  4M + 4M (2^22 each), B = 0.05 T tilted 30 deg from z to x,
                  seed 2, FP32 particle state (CELL32 storage), FP64 fields.
  electrode_1024  1024 x 1024 cells, centred 128 x 128-cell electrode at +100 V,
                  16M + 16M (2^24 each), B as above, seed 3, CELL32 storage.
  periodic_501x257  the reference's published run (Manya_Vansh_BMP.md:155-167):
                  500 x 256 cells (501 x 257 nodes) over 0.025 m x 0.0128 m
                  (dx = dy = 5e-5 m), periodic y, dt = 5e-12 s, electrons and
                  Xe+ ions (131.29 amu), Te = 10 eV, Ti = 0.5 eV, 9.6M of each,
                  200 V on the left wall and 0 V on the right and bottom walls
                  (a periodic geometry holds at most three rectangles,
                  particles.cpp:651-654; the bottom row bounds the top through
                  the wrap).  The published table gives no density or B: here
                  n0 = 1e16 m^-3 (Debye length 4.7 cells, omega_pe dt = 0.028)
                  and B = 0.  Seed 5, CELL32 storage.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

E_CHARGE = 1.602176634e-19
M_E = 9.1093837015e-31
AMU = 1.66053906660e-27
EPS = 8.854e-12          # the reference's M_PICDAT epsilon
N0 = 1e16
TE_EV = 2.0
TI_EV = 0.03
DX = 1e-4
M_ION = 40 * AMU

WALLS = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0], [0, 1, 1, 1, 0.0], [0, 0, 1, 0, 0.0]]


def plasma_dt(n0: float = N0, omega_dt: float = 0.1) -> float:
    omega_pe = math.sqrt(n0 * E_CHARGE ** 2 / (EPS * M_E))
    return omega_dt / omega_pe


@dataclass
class Workload:
    name: str
    NX: int
    NY: int
    n_per_species: int
    seed: int
    b_total: float
    b_tilt_deg: float
    electrode: tuple | None        # (x1, y1, x2, y2) fractions, or None
    electrode_v: float = 100.0
    layout: int = 0                # 0 F64, 1 CELL32
    steps: int = 100
    extra: dict = field(default_factory=dict)
    dx: float = DX                 # cell size (both directions)
    dt_s: float | None = None      # time step (default: omega_pe dt = 0.1 at N0)
    te_ev: float = TE_EV
    ti_ev: float = TI_EV
    ion_mass: float = M_ION
    periodic: int = 0              # periodic y
    walls: list | None = None      # boundary rectangles (default: all four walls at 0 V)

    @property
    def width(self):
        return self.NX * self.dx

    @property
    def height(self):
        return self.NY * self.dx

    @property
    def dt(self):
        return self.dt_s if self.dt_s is not None else plasma_dt()

    @property
    def bx(self):
        return self.b_total * math.sin(math.radians(self.b_tilt_deg))

    @property
    def bz(self):
        return self.b_total * math.cos(math.radians(self.b_tilt_deg))

    def rects(self):
        r = [list(w) for w in (self.walls if self.walls is not None else WALLS)]
        if self.electrode:
            x1, y1, x2, y2 = self.electrode
            r.append([x1, y1, x2, y2, self.electrode_v])
        return r

    def electrode_nodes(self):
        if not self.electrode:
            return None
        x1, y1, x2, y2 = self.electrode
        return [int(self.NX * x1), int(self.NY * y1), int(self.NX * x2), int(self.NY * y2)]

    def species(self):
        """(charge, mass, temperature eV) per species, electrons first."""
        return [(-E_CHARGE, M_E, self.te_ev), (E_CHARGE, self.ion_mass, self.ti_ev)]

    def scaled(self, NX=None, n=None, name=None):
        """Same physics and geometry fractions at another size."""
        w = Workload(**{**self.__dict__})
        if NX is not None:
            w.NX = w.NY = NX
        if n is not None:
            w.n_per_species = n
        w.name = name or f"{self.name}@{w.NX}x{w.n_per_species}"
        return w


ELECTRODE = (0.4375, 0.4375, 0.5625, 0.5625)

WORKLOADS = {
    "plain_cpu_128": Workload("plain_cpu_128", 128, 128, 100_000, 1, 0.01, 0.0, None, layout=0, steps=100),
    "electrode_512": Workload("electrode_512", 512, 512, 1 << 22, 2, 0.05, 30.0, ELECTRODE, layout=1, steps=500),
    "electrode_1024": Workload("electrode_1024", 1024, 1024, 1 << 24, 3, 0.05, 30.0, ELECTRODE, layout=1,
                               steps=1000),
    "periodic_501x257": Workload("periodic_501x257", 500, 256, 9_600_000, 5, 0.0, 0.0, None, layout=1, steps=10,
                                 dx=5e-5, dt_s=5e-12, te_ev=10.0, ti_ev=0.5, ion_mass=131.29 * AMU, periodic=1,
                                 walls=[[0, 0, 0, 1, 200.0], [1, 0, 1, 1, 0.0], [0, 0, 1, 0, 0.0]]),
}


def ensemble(nruns: int = 64, seed: int = 4, grids=(64, 128, 256), n_min: float = 5e4, n_max: float = 1e6):
    """Uneven ensemble: grids drawn from `grids`, particles per species
    log-uniform in [n_min, n_max], electrode voltage uniform in 10-200 V, Bz
    uniform in 0-0.1 T (no tilt), all drawn from `seed`; run k's particles use
    seed 1000 * seed + k."""
    import numpy as np
    rng = np.random.default_rng(seed)
    runs = []
    for k in range(nruns):
        NX = int(rng.choice(grids))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        v = float(rng.uniform(10.0, 200.0))
        bz = float(rng.uniform(0.0, 0.1))
        w = Workload(f"ens{k}", NX, NX, n, 1000 * seed + k, bz, 0.0, ELECTRODE, electrode_v=v, layout=1)
        runs.append(w)
    return runs


def generate(w: Workload, s: int):
    """Particles of species s of workload w from the library's generator."""
    import pic_hip
    q, m, t = w.species()[s]
    return pic_hip.generate_particles(w.seed, s, w.n_per_species, w.width, w.height, w.dx, w.dx, abs(q / m), t,
                                      excl=w.electrode_nodes())


def make_context(w: Workload, layout: int | None = None, particles=None, capacity_factor: float = 1.0,
                 **params):
    """Create a GPU context for workload w and upload its particles.

    `particles` may hold pre-generated (x, y, vx, vy, vz) per species;
    capacity_factor sizes the storage above the initial count (room for
    injected particles)."""
    import pic_hip
    lay = w.layout if layout is None else layout
    p = pic_hip.default_params(NX=w.NX, NY=w.NY, width=w.width, height=w.height, dt=w.dt, epsilon=EPS,
                               b0=w.bz, bx=w.bx, by=0.0, layout=lay, periodic=w.periodic, **params)
    ctx = pic_hip.Context(p)
    ctx.set_geometry(w.rects())
    for s, (q, m, _) in enumerate(w.species()):
        sid = ctx.add_species(q, m, w.n_per_species, N0, int(w.n_per_species * capacity_factor))
        parts = particles[s] if particles is not None else generate(w, s)
        ctx.upload(sid, *parts)
    return ctx
