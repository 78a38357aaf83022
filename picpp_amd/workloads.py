"""Workload definitions and seeded synthetic initial states.

Units are the reference's normalised Gaussian units with c = 1 (PIC++ keeps
c = 1, ``speed_of_light_normalized`` is commented out throughout, e.g.
PIC++/mover.cpp:234): lengths in electron skin depths d_e = c/omega_pe, times
in 1/omega_pe, masses in electron masses. With electron density n_e = 1 the
plasma frequency omega_pe^2 = 4 pi n_e q_e^2 / m_e = 1 fixes |q_e| = 1/sqrt(4 pi).
Particle weight is the number of physical particles a macro-particle carries,
n V_cell / ppc, as in the reference's weight = concentration/ppc * volume
(PIC++/mover.cpp:935-937 uses the same product for injected particles).

The grid is the reference's ghosted grid: n physical cells per axis, two ghost
layers each side (N = n + 4), node 0 at coordinate L - 2 dx with L = n dx
(PIC++/initialization.cpp:71-202), so owned particles sit in [L, 2L).

Initial states are synthetic: the reference seeds rand() with time(NULL)
(PIC++/random.cpp:35, main.cpp:73), so its states cannot be reproduced; every
state here is generated from an explicit seed.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

Q_E = 1.0 / math.sqrt(4.0 * math.pi)  # |electron charge| in these units

# PIC++ ParticleTypes (PIC++/particleTypes.h:4)
ELECTRON, PROTON, POSITRON = 0, 1, 2


@dataclasses.dataclass
class Species:
    name: str
    ptype: int            # reference ParticleTypes index
    mass: float           # in electron masses
    charge: float         # signed, in these units
    ppc: int              # macro-particles per cell
    temperature: float    # kT / (m c^2) of this species in its rest frame
    density: float = 1.0  # physical density in n_e units
    drift_gamma: float = 1.0  # bulk Lorentz factor of the drift
    drift_axis: int = 2
    drift_sign: int = 1

    @property
    def weight_per_volume(self) -> float:
        return self.density / self.ppc


@dataclasses.dataclass
class Workload:
    name: str
    n: Tuple[int, int, int]
    dx: float
    dt: float
    steps: int
    species: List[Species]
    b0: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    e0: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    seed: int = 42
    gmres_tolerance: float = 1e-8
    # x boundary: "periodic", or "wall" (a perfectly conducting reflecting wall at the left
    # end and a constant-field free boundary at the right end, PIC++/initialization.cpp:5201-5210)
    boundary: str = "periodic"
    e_right: Tuple[float, float, float] = (0.0, 0.0, 0.0)  # fields held at the free right boundary
    b_right: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    # wall boundary only: inject the flow at the right end each step and replace particles that
    # leave there (PIC++/mover.cpp:910-1000, initialization.cpp:8281-8338)
    inject: bool = False
    # smoothing of rhohat, J and the dielectric tensor after the moments (smoothingCount passes
    # of the 27-point kernel with weights (s/2, 1-s, s/2) per axis, PIC++/smoothing.cpp)
    smoothing_count: int = 0
    smoothing_parameter: float = 0.5
    # divergence cleanup after the push: 1 newE (PIC++/divergencecleanup.cpp:20-392), 2 newB
    # (:939-1306), 3 both; and the cut wave number of the Fourier filtering of newE, newB
    # (PIC++/fields_filtering.cpp:21-295; 0: off)
    divergence_cleanup: int = 0
    filter_cut_wave_number: int = 0
    run_kT: float = 0.0  # kB T of the run (Simulation::temperature) in m c^2 units: the Juttner branch choice

    @property
    def N(self) -> Tuple[int, int, int]:
        """Ghosted extent per axis: n + 4, but 1 along a one-cell z (the reference's 2-D XY
        grids have no ghost layers there, PIC++/initialization.cpp:170-174)."""
        return tuple(1 if (a == 2 and v == 1) else v + 4 for a, v in enumerate(self.n))

    @property
    def cell_volume(self) -> float:
        return self.dx ** 3

    @property
    def particles_total(self) -> int:
        return sum(s.ppc for s in self.species) * int(np.prod(self.n))

    def capacity_needed(self, steps: Optional[int] = None) -> int:
        """Particle capacity for `steps` steps (default: the workload's): the initial particles
        plus, with injection, the slabs injected at the right end (|V0_x| dt / dx ppc of them
        per step and species, one particle per (y, z) column each)."""
        total = self.particles_total
        if self.inject:
            steps = self.steps if steps is None else steps
            cols = self.n[1] * self.n[2]
            slabs = abs(self.flow[0]) * self.dt * steps / self.dx + 1
            total += sum(int(math.ceil(slabs * s.ppc)) * cols for s in self.species)
        return total

    def weight(self, s: Species) -> float:
        return s.density * self.cell_volume / s.ppc

    @property
    def flow(self) -> Tuple[float, float, float]:
        """Flow velocity V0 of the injected plasma: the drift of the first species."""
        s = self.species[0]
        v = [0.0, 0.0, 0.0]
        if s.drift_gamma > 1.0:
            v[s.drift_axis] = s.drift_sign * math.sqrt(1.0 - 1.0 / (s.drift_gamma * s.drift_gamma))
        return tuple(v)

    def origin(self) -> Tuple[float, float, float]:
        """Coordinate of node 0 per axis: L - 2 dx, or L along a one-cell z
        (PIC++/initialization.cpp:115-131, :194)."""
        return tuple(v * self.dx - (0 if (a == 2 and v == 1) else 2) * self.dx for a, v in enumerate(self.n))


def electron_proton(ppc: int, theta_e: float, mass_ratio: float, theta_p: Optional[float] = None) -> List[Species]:
    """Electron-proton plasma at a common temperature unless theta_p is given."""
    if theta_p is None:
        theta_p = theta_e / mass_ratio  # same kT
    return [
        Species("electron", ELECTRON, 1.0, -Q_E, ppc, theta_e),
        Species("proton", PROTON, mass_ratio, Q_E, ppc, theta_p),
    ]


def b_for_gyro(omega_ratio: float) -> float:
    """|B| that gives omega_ce / omega_pe = omega_ratio for m_e = 1, |q_e| = Q_E."""
    return omega_ratio / Q_E


def config1(n=(32, 32, 8), ppc=20, steps=50, seed=42) -> Workload:
    """Plain FP64 implicit step, the reference's CPU-runnable case: periodic 32x32x8,
    dx = 1 d_e, e-/p+ with m_p/m_e = 100, 20 ppc per species, kT/mc^2 = 0.01,
    omega_ce/omega_pe = 0.5 along x, dt = 0.1/omega_pe, 50 steps, seed 42."""
    return Workload("implicit-e-p-%dx%dx%d-ppc%d" % (n[0], n[1], n[2], ppc), tuple(n), 1.0, 0.1, steps,
                    electron_proton(ppc, 0.01, 100.0), b0=(b_for_gyro(0.5), 0.0, 0.0), seed=seed)


def config2(n=(128, 128, 128), ppc=64, steps=200, seed=42) -> Workload:
    """Same physics at one-rank production size: 128^3 cells, 64 ppc per species."""
    return Workload("implicit-e-p-%dx%dx%d-ppc%d" % (n[0], n[1], n[2], ppc), tuple(n), 1.0, 0.1, steps,
                    electron_proton(ppc, 0.01, 100.0), b0=(b_for_gyro(0.5), 0.0, 0.0), seed=seed)


def mini(n=(6, 5, 5), ppc=4, steps=5, seed=7) -> Workload:
    """Small case for golden vectors and fast parity tests."""
    return Workload("mini-%dx%dx%d-ppc%d" % (n[0], n[1], n[2], ppc), tuple(n), 1.0, 0.1, steps,
                    electron_proton(ppc, 0.01, 100.0), b0=(b_for_gyro(0.5), 0.0, 0.0), seed=seed)


def shock(n=(1024, 64, 64), ppc=16, gamma0=10.0, sigma=0.1, theta=1e-4, mass_ratio=64.0, dt=0.1, steps=500, seed=11,
          name="shock", inject=True) -> Workload:
    """Relativistic shock setup: a cold e-/p+ flow with bulk Lorentz factor gamma0 moving along
    -x onto the reflecting wall at the left end, transverse B0 along z with magnetisation
    sigma = B0^2 / (4 pi gamma0 n (m_e + m_p) c^2), and the motional field E0 = -V0 x B0
    (PIC++/initialization.cpp:5201-5210). The right boundary holds E0, B0; with `inject` the
    flow is injected there every step (PIC++/simulation.cpp:281-289)."""
    v = math.sqrt(1.0 - 1.0 / (gamma0 * gamma0))
    b = math.sqrt(4.0 * math.pi * sigma * gamma0 * (1.0 + mass_ratio))
    e0 = (0.0, -v * b, 0.0)  # -V0 x B0 with V0 = (-v, 0, 0), B0 = (0, 0, b)
    sp = electron_proton(ppc, theta, mass_ratio)
    for s in sp:
        s.drift_gamma, s.drift_axis, s.drift_sign = gamma0, 0, -1
    return Workload("%s-%dx%dx%d-ppc%d" % (name, n[0], n[1], n[2], ppc), tuple(n), 1.0, dt, steps, sp, b0=(0.0, 0.0, b),
                    e0=e0, seed=seed, boundary="wall", e_right=e0, b_right=(0.0, 0.0, b), inject=inject)


def weibel(n=(192, 192, 192), ppc=32, steps=300, seed=7) -> Workload:
    """Counterstreaming e-/e+ beams along +/-z, drift gamma 5, kT/mc^2 = 0.01 in the beam frame."""
    sp = []
    for name, ptype, q in (("electron", ELECTRON, -Q_E), ("positron", POSITRON, Q_E)):
        for sign in (1, -1):
            sp.append(Species("%s%+d" % (name, sign), ptype, 1.0, q, ppc, 0.01, density=0.5,
                              drift_gamma=5.0, drift_axis=2, drift_sign=sign))
    return Workload("weibel-%d^3-ppc%d" % (n[0], ppc), tuple(n), 1.0, 0.1, steps, sp, seed=seed)


# ------------------------------------------------------------------ sampling
def _mj_table(theta: float, npts: int = 65536):
    """Tabulated CDF of |u| for the Maxwell-Juettner distribution
    f(u) du ~ u^2 exp(-(gamma - 1)/theta) du, gamma = sqrt(1 + u^2)."""
    gmax = 1.0 + 60.0 * theta
    umax = math.sqrt(gmax * gmax - 1.0)
    u = np.linspace(0.0, umax, npts)
    g = np.sqrt(1.0 + u * u)
    # (gamma - 1) in the compensated form u^2/(gamma+1) keeps small-theta tails exact
    f = u * u * np.exp(-(u * u / (g + 1.0)) / theta)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(u))])
    cdf /= cdf[-1]
    return u, cdf


def sample_maxwell_juettner(rng: np.random.Generator, theta: float, count: int) -> np.ndarray:
    """Isotropic Maxwell-Juettner momenta u = p/(m c), shape (count, 3)."""
    ut, cdf = _mj_table(theta)
    mag = np.interp(rng.random(count), cdf, ut)
    cost = 2.0 * rng.random(count) - 1.0
    phi = 2.0 * math.pi * rng.random(count)
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost * cost))
    return np.stack([mag * sint * np.cos(phi), mag * sint * np.sin(phi), mag * cost], axis=1)


def boost(rng: np.random.Generator, u: np.ndarray, gamma_d: float, axis: int, sign: int) -> np.ndarray:
    """Boost rest-frame momenta along `axis` with the volume-transform flipping
    method (Zenitani 2015, Phys. Plasmas 22, 042116)."""
    if gamma_d <= 1.0:
        return u
    beta = sign * math.sqrt(1.0 - 1.0 / (gamma_d * gamma_d))
    g = np.sqrt(1.0 + np.sum(u * u, axis=1))
    ua = u[:, axis].copy()
    flip = (-beta * ua / g) > rng.random(len(ua))
    ua[flip] = -ua[flip]
    out = u.copy()
    out[:, axis] = gamma_d * (ua + beta * g)
    return out


@dataclasses.dataclass
class InitialState:
    species: np.ndarray    # int32 (np,)
    ids: np.ndarray        # int32 (np,)
    x: np.ndarray          # float64 (np, 3) absolute coordinates
    p: np.ndarray          # float64 (np, 3) momentum m*gamma*v
    E: np.ndarray          # float64 (N+1, N+1, N+1, 3) ghosted nodes
    B: np.ndarray          # float64 (N, N, N, 3) ghosted cells


def periodic_fill_nodes(owned: np.ndarray) -> np.ndarray:
    """Ghosted node array (N+1 per axis) from owned nodes 2..n+1, periodic."""
    n = owned.shape[:3]
    idx = [np.mod(np.arange(v + 5) - 2, v) for v in n]
    return owned[np.ix_(idx[0], idx[1], idx[2])]


def periodic_fill_cells(owned: np.ndarray) -> np.ndarray:
    """Ghosted cell array (N per axis) from owned cells 2..n+1, periodic."""
    n = owned.shape[:3]
    idx = [np.mod(np.arange(v + 4) - 2, v) for v in n]
    return owned[np.ix_(idx[0], idx[1], idx[2])]


def make_initial_state(w: Workload, perturb: float = 0.0) -> InitialState:
    """Seeded initial state: positions uniform within each cell, momenta
    Maxwell-Juettner (boosted for drifting species), uniform E0/B0 plus an optional
    smooth perturbation of relative size `perturb` (used by parity tests so that
    gathers see non-uniform fields)."""
    rng = np.random.default_rng(w.seed)
    n = np.array(w.n)
    ncell = int(np.prod(n))
    lo = n * w.dx  # first owned coordinate per axis
    sps, xs, ps = [], [], []
    cells = np.stack(np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij"), -1).reshape(-1, 3)
    for si, s in enumerate(w.species):
        cnt = ncell * s.ppc
        c = np.repeat(cells, s.ppc, axis=0)
        x = lo + (c + rng.random((cnt, 3))) * w.dx
        u = sample_maxwell_juettner(rng, s.temperature, cnt)
        u = boost(rng, u, s.drift_gamma, s.drift_axis, s.drift_sign)
        sps.append(np.full(cnt, si, np.int32))
        xs.append(x)
        ps.append(u * s.mass)
    sp = np.concatenate(sps)
    x = np.concatenate(xs)
    p = np.concatenate(ps)
    # fields on owned nodes (node i at lo + i dx) and cells (centre lo + (i + 1/2) dx)
    Eo = np.zeros(tuple(n) + (3,))
    Bo = np.zeros(tuple(n) + (3,))
    Eo[...] = np.array(w.e0)
    Bo[...] = np.array(w.b0)
    if perturb:
        grid_n = np.meshgrid(*[np.arange(v) / v for v in n], indexing="ij")
        grid_c = np.meshgrid(*[(np.arange(v) + 0.5) / v for v in n], indexing="ij")
        b0 = max(1.0, float(np.linalg.norm(w.b0)))
        for c in range(3):
            ph = 2 * math.pi * (grid_n[0] * (c + 1) + grid_n[1] * 1 + grid_n[2] * (2 - c % 2))
            Eo[..., c] += perturb * b0 * np.sin(ph + 0.3 * c)
            ph = 2 * math.pi * (grid_c[0] * 1 + grid_c[1] * (c % 2 + 1) + grid_c[2] * 1)
            Bo[..., c] += perturb * b0 * np.cos(ph + 0.7 * c)
    E = periodic_fill_nodes(Eo)
    if w.boundary == "wall":
        # tangential E vanishes on the conducting wall and in the ghost layers behind it
        # (PIC++/initialization.cpp:5270-5283)
        E[:3, :, :, 1:] = 0.0
    return InitialState(sp, np.arange(len(sp), dtype=np.int32), x, p, E, periodic_fill_cells(Bo))


YEE_NAMES = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def yee_dims(w: Workload, name: str):
    """Shape of a ghosted Yee (Buneman) array of the explicit solver, as the reference
    allocates it (PIC++/initialization.cpp:7146-7174): E and J component c have N points along
    axis c and N + 1 along the others; B component c the opposite."""
    N = w.N
    c = "xyz".index(name[1])
    b = name[0] == "B"
    return tuple(N[a] if (a == c) != b else N[a] + 1 for a in range(3))


def yee_initial_fields(w: Workload, perturb: float = 0.0):
    """Yee fields of the explicit solver: the uniform E0, B0 plus an optional smooth periodic
    perturbation of relative size `perturb`, on every point of the ghosted arrays (each point
    at its staggered position: a component's own axis at cell centres for E, nodes for B)."""
    out = {}
    b0 = max(1.0, float(np.linalg.norm(w.b0)))
    for name in YEE_NAMES:
        dims = yee_dims(w, name)
        c = "xyz".index(name[1])
        is_b = name[0] == "B"
        base = (w.b0 if is_b else w.e0)[c]
        a = np.full(dims, float(base))
        if perturb:
            # coordinates in units of the box, ghost index 2 at 0; cell-centred axes shifted by 1/2
            cc = []
            for ax in range(3):
                half = 0.5 if (ax == c) != is_b else 0.0
                cc.append((np.arange(dims[ax]) - 2 + half) / w.n[ax])
            g = np.meshgrid(*cc, indexing="ij")
            if is_b:
                ph = 2 * math.pi * (g[0] * 1 + g[1] * (c % 2 + 1) + g[2] * 1)
                a += perturb * b0 * np.cos(ph + 0.7 * c)
            else:
                ph = 2 * math.pi * (g[0] * (c + 1) + g[1] * 1 + g[2] * (2 - c % 2))
                a += perturb * b0 * np.sin(ph + 0.3 * c)
        out[name] = a
    return out


def species_arrays(w: Workload):
    """(mass, charge, weight, ptype) arrays per species."""
    return (np.array([s.mass for s in w.species], np.float64),
            np.array([s.charge for s in w.species], np.float64),
            np.array([w.weight(s) for s in w.species], np.float64),
            np.array([s.ptype for s in w.species], np.int32))


def make_state_on_device(w, torch, device):
    """Seeded synthetic state on the GPU: uniform positions in each cell, Maxwell-Juettner
    momenta (tabulated inverse CDF, as picpp_amd.workloads does on the host)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(w.seed)
    n = w.n
    ncell = n[0] * n[1] * n[2]
    xs, ps, counts = [], [], []
    for s in w.species:
        cnt = ncell * s.ppc
        cid = torch.arange(ncell, device=device, dtype=torch.int64).repeat_interleave(s.ppc)
        ci = torch.stack([cid // (n[1] * n[2]), (cid // n[2]) % n[1], cid % n[2]]).to(torch.float64)
        del cid
        lo = torch.tensor([v * w.dx for v in n], device=device, dtype=torch.float64)[:, None]
        x = lo + (ci + torch.rand((3, cnt), generator=gen, device=device, dtype=torch.float64)) * w.dx
        del ci
        ut, cdf = _mj_table(s.temperature)
        ut = torch.tensor(ut, device=device)
        cdf = torch.tensor(cdf, device=device)
        r = torch.rand(cnt, generator=gen, device=device, dtype=torch.float64)
        k = torch.searchsorted(cdf, r).clamp(1, len(cdf) - 1)
        t = ((r - cdf[k - 1]) / (cdf[k] - cdf[k - 1]).clamp_min(1e-300)).clamp(0, 1)
        mag = ut[k - 1] + t * (ut[k] - ut[k - 1])
        del r, k, t
        cost = 2 * torch.rand(cnt, generator=gen, device=device, dtype=torch.float64) - 1
        phi = 2 * math.pi * torch.rand(cnt, generator=gen, device=device, dtype=torch.float64)
        sint = torch.sqrt((1 - cost * cost).clamp_min(0))
        u = torch.stack([mag * sint * torch.cos(phi), mag * sint * torch.sin(phi), mag * cost])
        del cost, phi, sint, mag
        if s.drift_gamma > 1.0:  # boost along the drift axis (volume-transform flipping, as workloads.boost)
            beta = s.drift_sign * math.sqrt(1.0 - 1.0 / (s.drift_gamma * s.drift_gamma))
            g = torch.sqrt(1.0 + (u * u).sum(0))
            ua = u[s.drift_axis]
            flip = (-beta * ua / g) > torch.rand(cnt, generator=gen, device=device, dtype=torch.float64)
            ua = torch.where(flip, -ua, ua)
            u[s.drift_axis] = s.drift_gamma * (ua + beta * g)
            del g, ua, flip
        p = u * s.mass
        del u
        xs.append(x)
        ps.append(p)
        counts.append(cnt)
    x = torch.cat(xs, 1).contiguous()
    p = torch.cat(ps, 1).contiguous()
    return x, p, counts
