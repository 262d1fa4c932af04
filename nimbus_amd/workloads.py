"""Named workloads: grids, large-step settings and seeded initial conditions.

c0  64x64x40, dx = dy = 1000 m, dz = 500 m (20 km top), dt = 6 s, 6 substeps, 10 steps
c1  160x160x80, dx = dy = 500 m, dz = 250 m, dt = 3 s, 6 substeps, 1 simulated hour
c2  400x400x100, dx = dy = 250 m, dz = 200 m, dt = 1.5 s, 8 substeps, 1 simulated hour,
    three warm bubbles 15 km apart
c3  32-member ensemble of uneven grids drawn from seed 3 (nx = ny in {48, 64, 96, 128, 160},
    nz in {40, 60, 80}, dx in {500, 1000} m, dt = long_run_dt(dx, dz), 6 substeps, 30
    simulated minutes) with per-member sounding perturbations
All use the analytic Weisman-Klemp-type sounding with a quarter-circle
hodograph, +2 K bubbles of 10 km x 1.4 km radius centred 1.4 km above ground,
and seeded uniform +-0.1 K theta noise. Unless a workload sets sponge=False
or sgs=False, its large steps carry the Rayleigh sponges (upper sponge from
12 km, 8-cell w absorber, 10 km lateral band; lib.sponge_preset) and the
horizontal Smagorinsky closure (lib.sgs_preset) of the reference Kessler
presets.
"""
import dataclasses

import numpy as np

from . import lib as L


@dataclasses.dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dz: float
    dt: float
    substeps: int
    steps: int
    bubbles: int = 1
    seed: int = 0
    member: int = 0
    sounding: dict = dataclasses.field(default_factory=dict)
    sponge: bool = True
    sgs: bool = True

    @property
    def grid(self):
        return L.make_grid(self.nx, self.ny, self.nz, self.dx, self.dx, self.dz)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    def config(self, acoustic_mode=L.ACOUSTIC_EXPLICIT, **kw):
        if self.sponge:
            kw.setdefault("sponge", L.sponge_preset(self.dx))
        if self.sgs:
            kw.setdefault("sgs", L.sgs_preset())
        return L.step_config(self.dt, self.substeps, acoustic_mode=acoustic_mode, **kw)

    def sounding_struct(self):
        return L.default_sounding(**self.sounding)

    def base(self):
        return L.base_state(self.grid, self.sounding_struct())

    def ic_params(self):
        g = self.grid
        xc = (self.nx - 2) * self.dx / 2.0
        yc = (self.ny - 2) * self.dx / 2.0
        if self.bubbles == 3:
            bubbles = [(xc + off, yc, 1400.0, 10000.0, 1400.0, 2.0) for off in (-15000.0, 0.0, 15000.0)]
        else:
            bubbles = [(xc, yc, 1400.0, 10000.0, 1400.0, 2.0)]
        return L.ic_params(g, self.sounding_struct(), bubbles=bubbles, noise_amp=0.1, seed=self.seed,
                           member=self.member)

    def initial_state(self, dtype=np.float32):
        return L.initial_state(self.grid, self.ic_params(), dtype)

    def describe(self):
        return (f"{self.nx}x{self.ny}x{self.nz} dx={self.dx:g}m dz={self.dz:g}m dt={self.dt:g}s "
                f"substeps={self.substeps}")


C0 = Workload("c0", 64, 64, 40, 1000.0, 500.0, 6.0, 6, 10)
C1 = Workload("c1", 160, 160, 80, 500.0, 250.0, 3.0, 6, 1200)
C2 = Workload("c2", 400, 400, 100, 250.0, 200.0, 1.5, 8, 2400, bubbles=3)
NAMED = {"c0": C0, "c1": C1, "c2": C2}


def _splitmix(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


class SeededStream:
    """SplitMix64 stream of uniform [0, 1) draws."""

    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def uniform(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        return (_splitmix(self.state) >> 11) / float(1 << 53)

    def choice(self, values):
        return values[min(int(self.uniform() * len(values)), len(values) - 1)]

    def between(self, lo, hi):
        return lo + (hi - lo) * self.uniform()


def long_run_dt(dx, dz):
    """Large step for the generated long runs (ensemble members, serving
    requests): 6 s per km of the finer of dx and dz. The momentum advection
    of the slow step (Heun RK2 with fifth-order upwind differences) is stable
    only up to Courant numbers of a few tenths (linear analysis of Wicker and
    Skamarock 2002, Mon. Wea. Rev. 130), so with updrafts of up to 50 m/s the
    vertical step is kept at w dt / dz <= 0.3, i.e. dt <= 6 dz / 1000 s.
    With 6 dx / 1000 s alone, a 90-minute request at dz = 500 m, dt = 5.2 s
    blew up in both precisions after about 70 minutes."""
    return 6.0 * min(dx, dz) / 1000.0


def perturbed_sounding(stream):
    """Per-member sounding perturbation: surface theta +-2 K, surface qv 11-16 g/kg,
    shear 10-40 m/s, hodograph curvature 0-1."""
    return {"theta_sfc": 300.0 + stream.between(-2.0, 2.0), "qv_cap": stream.between(0.011, 0.016),
            "shear": stream.between(10.0, 40.0), "curvature": stream.uniform()}


def ensemble_c3(n_members=32, seed=3):
    """The 32-member uneven-grid ensemble (30 simulated minutes per member)."""
    s = SeededStream(seed)
    members = []
    for m in range(n_members):
        n = s.choice([48, 64, 96, 128, 160])
        nz = s.choice([40, 60, 80])
        dx = s.choice([500.0, 1000.0])
        dt = long_run_dt(dx, 20000.0 / nz)
        members.append(Workload(f"c3m{m}", n, n, nz, dx, 20000.0 / nz, dt, 6, int(round(1800.0 / dt)),
                                seed=seed, member=m, sounding=perturbed_sounding(s)))
    return members
