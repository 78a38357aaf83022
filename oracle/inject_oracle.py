"""CPU restatement of the injection at the free right end and of the replacement of
escaped particles: test infrastructure only (tests/ import it as the checker of the HIP
injection kernels; the product never does).

Reference: Simulation::injectNewParticles (PIC++/mover.cpp:910-1000), the injectionLength
update of the step loop (PIC++/simulation.cpp:125-127, :281-283), addToPreserveChargeGlobal
(PIC++/initialization.cpp:8281-8338), createParticle's Maxwellian and isotropic Juttner branches
(PIC++/initialization.cpp:8479-8535) with the tabulated CDF of evaluateJuttnerFunction
(:7372-7394) and the samplers of PIC++/random.cpp:43-66, Particle::getVelocity / addVelocity / setMomentumByV
(PIC++/particle.cpp:216-232, :291-331, :337-364).

The reference draws from rand(); the HIP path and this restatement draw from Philox4x64-10
with the counters listed in picpp_amd/csrc/pic_inject.hip, so the two agree sample by
sample. What pins each part:
- philox4x64_10: the Random123 known-answer vectors and numpy.random.Philox, an independent
  implementation of the same generator (tests/test_inject_oracle.py);
- plan (slab counts and x positions): the compiled reference's own injection
  (tests/golden/inject_ref.npz), exactly;
- the placement of slab particles and replacements (columns, common in-cell offset of a
  slab, x of replacements, kept y and z): the same golden vectors;
- flow_momentum / boost: closed forms -- with V0 = 0 the momenta are sqrt(theta) m N(0, 1) per
  axis; the boost preserves the invariant mass (gamma^2 - u^2 = 1) and maps the rest-frame mean
  to gamma_V V along V0; with theta = 0 every particle gets exactly gamma_V m V0;
- the Juttner branches (rest_u from theta = 0.1 on): the mean Lorentz factor of the samples
  against the Maxwell-Juttner closed form <gamma> = K1(1/theta) / K2(1/theta) + 3 theta (tabulated
  branch) and the sampler's own target (rejection branch), isotropy, and the compiled
  reference's own hot injection (the means and spreads of its momenta).
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np

MASK = (1 << 64) - 1
M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157  # Philox4x64 multipliers
W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B  # Weyl key increments
REL_PRECISION = 0.0001  # relativisticPrecision, PIC++/constants.h:39
TWO_PI = 2.0 * (4.0 * math.atan2(1.0, 1.0))


def philox4x64_10(ctr: Sequence[int], key: Sequence[int]) -> List[int]:
    """Philox4x64 with 10 rounds (Salmon et al., SC 2011): 4 x 64-bit outputs."""
    c = [int(v) & MASK for v in ctr]
    k0, k1 = int(key[0]) & MASK, int(key[1]) & MASK
    for r in range(10):
        if r:
            k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [(p1 >> 64) ^ c[1] ^ k0, p1 & MASK, (p0 >> 64) ^ c[3] ^ k1, p0 & MASK]
    return c


def uniform53(w: int) -> float:
    return float(w >> 11) * 2.0 ** -53


def box_muller(wa: int, wb: int) -> Tuple[float, float]:
    r = math.sqrt(-2.0 * math.log(1.0 - uniform53(wa)))
    a = TWO_PI * uniform53(wb)
    return r * math.cos(a), r * math.sin(a)


JUTTNER_POINTS = 4096  # randomParameter, PIC++/constants.h:8


def juttner_table(theta: float) -> Tuple[np.ndarray, np.ndarray]:
    """evaluateJuttnerFunction (PIC++/initialization.cpp:7372-7394): u_i = i 20 theta / (points - 1)
    and the running sum of u^2 exp(-sqrt(1 + u^2) / theta) du / 10 over 10 sub-steps per interval,
    normalised to end at exactly 1."""
    u = np.zeros(JUTTNER_POINTS)
    F = np.zeros(JUTTNER_POINTS)
    du = 20 * theta / (JUTTNER_POINTS - 1)
    sdu = du / 10
    for i in range(1, JUTTNER_POINTS):
        u[i] = u[i - 1] + du
        f = F[i - 1]
        lu = u[i - 1]
        for _ in range(10):
            f += math.exp(-math.sqrt(1 + lu * lu) / theta) * lu * lu * sdu
            lu += sdu
        F[i] = f
    F /= F[-1]
    F[-1] = 1.0
    return u, F


def juttner_inverse(table: Tuple[np.ndarray, np.ndarray], y: float) -> float:
    """dichotomySolver (PIC++/specialmath.cpp:931-947) on the table: halve [0, points - 1] until
    one step wide, then interpolate linearly."""
    u, F = table
    lo, hi = 0, JUTTNER_POINTS - 1
    while lo + 1 != hi:
        mid = (lo + hi) // 2
        if F[mid] > y:
            hi = mid
        elif F[mid] < y:
            lo = mid
        else:
            return float(u[mid])
    return float((u[hi] - u[lo]) * (y - F[lo]) / (F[hi] - F[lo]) + u[lo])


def rest_u(theta: float, theta_run: float, table, ctr: Sequence[int], key: Sequence[int]) -> List[float]:
    """createParticle's rest-frame u = p / (m c) (PIC++/initialization.cpp:8505-8535): below
    theta = 0.1 u_i = sqrt(theta) N(0, 1); from 0.1 on an isotropic direction and |u| from the
    table (theta_run < 2, maxwellJuttnerMomentumColdDistribution, PIC++/random.cpp:43-48) or from
    the rejection sampler (maxwellJuttnerMomentumHotDistribution, :50-66, U1 kept over the
    attempts), with the counters of picpp_amd/csrc/pic_inject.hip."""
    w = philox4x64_10(ctr, key)
    if theta < 0.1:
        n0, n1 = box_muller(w[0], w[1])
        n2, _ = box_muller(w[2], w[3])
        st = math.sqrt(theta)
        return [st * n0, st * n1, st * n2]
    if theta_run < 2:
        a = juttner_inverse(table, uniform53(w[0]))
    else:
        x1 = uniform53(w[0])
        a = -theta_run * math.log(uniform53(w[1]) * uniform53(w[2]) * uniform53(w[3]))
        att = 1
        while att < 64 and math.exp((a - math.sqrt(1 + a * a)) / theta_run) < x1:
            w = philox4x64_10((ctr[0], ctr[1], ctr[2], att), key)
            a = -theta_run * math.log(uniform53(w[1]) * uniform53(w[2]) * uniform53(w[3]))
            att += 1
    d = philox4x64_10((ctr[0], ctr[1], ctr[2], 1 << 40), key)
    uz = a * (2 * uniform53(d[0]) - 1)
    phi = TWO_PI * uniform53(d[1])
    un = math.sqrt(a * a - uz * uz)
    return [un * math.cos(phi), un * math.sin(phi), uz]


def flow_momentum(flow: Sequence[float], m: float, theta: float, w: Sequence[int]) -> np.ndarray:
    """p = m gamma v of a Maxwellian particle (p_i = sqrt(m k T) N(0,1),
    PIC++/initialization.cpp:8507-8514) from the generator words w, boosted by V0."""
    n0, n1 = box_muller(w[0], w[1])
    n2, _ = box_muller(w[2], w[3])
    st = math.sqrt(theta)
    return boost(flow, m, [st * n0, st * n1, st * n2])


def boost(flow: Sequence[float], m: float, u: Sequence[float]) -> np.ndarray:
    """The rest-frame u boosted by V0 with the relativistic velocity addition
    (PIC++/particle.cpp:291-331), p from v as setMomentumByV (particle.cpp:337-364)."""
    g1 = math.sqrt((u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) + 1.0)
    v = [u[0] / g1, u[1] / g1, u[2] / g1]
    V2 = flow[0] * flow[0] + flow[1] * flow[1] + flow[2] * flow[2]
    if V2 > 0.0:
        V = math.sqrt(V2)
        e = [flow[0] / V, flow[1] / V, flow[2] / V]
        gV = 1.0 / math.sqrt(1.0 - V2)
        vpar = v[0] * e[0] + v[1] * e[1] + v[2] * e[2]
        den = 1.0 + V * vpar
        npar = (V + vpar) / den
        v = [npar * e[c] + (v[c] - vpar * e[c]) / (gV * den) for c in range(3)]
    v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    f = m if v2 < REL_PRECISION else m * (1.0 / math.sqrt(1.0 - v2))
    return np.array([v[0] * f, v[1] * f, v[2] * f])


class Injector:
    """Per-run injection state: injectionLength per species, step and slab numbers, next id.

    `w` is a picpp_amd.workloads.Workload with boundary "wall" and inject on; `next_id` is one
    past the largest uploaded id."""

    def __init__(self, w, next_id: int):
        self.w = w
        self.n = w.n
        self.d = w.dx
        self.x0 = w.origin()
        self.ppc = [s.ppc for s in w.species]
        self.theta = [s.temperature for s in w.species]
        self.mass = [s.mass for s in w.species]
        self.theta_run = [getattr(w, "run_kT", 0.0) / s.mass for s in w.species]
        self.tables = [juttner_table(t) if t >= 0.1 and p > 0 else None for t, p in zip(self.theta, [s.ppc for s in w.species])]
        self.flow = w.flow
        self.seed = w.seed
        # injectionLength at the loop start: particesDeltaX - 1e-4 dx, particesDeltaX = dx / ppc
        # (PIC++/simulation.cpp:125-127, initialization.cpp:7341)
        self.len = [self.d / p - 0.0001 * self.d if p > 0 else 0.0 for p in self.ppc]
        self.step = 0
        self.slabs = 0
        self.next_id = next_id

    def right_edge(self) -> float:
        return self.x0[0] + (self.n[0] + 4 - 2) * self.d  # xgrid[xnumberAdded - 1 - additionalBinNumber]

    def plan(self) -> List[Tuple[int, float, int]]:
        """(species, x, slab number) of this step's slabs (PIC++/simulation.cpp:281-283,
        mover.cpp:910-925)."""
        out = []
        xr = self.right_edge()
        for s in range(len(self.ppc)):
            self.len[s] += abs(self.flow[0] * self.w.dt)
            p = self.ppc[s]
            if p > 0 and p * self.len[s] >= self.d:
                cnt = int(self.len[s] * p / self.d)
                for _ in range(cnt):
                    self.len[s] -= self.d / p
                    out.append((s, xr - self.len[s], self.slabs))
                    self.slabs += 1
        return out

    def step_particles(self, escaped: Sequence[Tuple[int, float, float]]) -> Dict[str, np.ndarray]:
        """New particles of one step: the injected slabs (PIC++/mover.cpp:940-998), then one
        replacement per escaped particle (species, y, z), given in ascending index order
        (PIC++/initialization.cpp:8286-8317). Returns species, ids, x [n, 3], p [n, 3] and
        kind (0 injected, 1 replacement)."""
        n1, n2 = self.n[1], self.n[2]
        sp, ids, xs, ps, kind = [], [], [], [], []
        key = (self.seed, 0)
        for s, x, slab in self.plan():
            w = philox4x64_10((self.step, (2 << 32) | slab, 0, 0), key)
            dy, dz = self.d * uniform53(w[0]), self.d * uniform53(w[1])
            for col in range(n1 * n2):
                j, k = 2 + col // n2, 2 + col % n2
                y = (self.x0[1] + j * self.d) + dy
                z = (self.x0[2] + k * self.d) + dz
                u = rest_u(self.theta[s], self.theta_run[s], self.tables[s], (self.step, (0 << 32) | slab, col, 0), key)
                sp.append(s)
                ids.append(self.next_id)
                self.next_id += 1
                xs.append((x, y, z))
                ps.append(boost(self.flow, self.mass[s], u))
                kind.append(0)
        xr = self.right_edge()
        for r, (s, y, z) in enumerate(escaped):
            wy = philox4x64_10((self.step, (1 << 32) | r, 1, 0), key)
            pos = [xr, y, z]
            for a in (1, 2):
                lo, hi = self.x0[a] + 2 * self.d, self.x0[a] + (self.n[a] + 4 - 2) * self.d
                if pos[a] < lo:
                    pos[a] = lo + self.d * uniform53(wy[2 * (a - 1)])
                if pos[a] > hi:
                    pos[a] = hi - self.d * uniform53(wy[2 * (a - 1) + 1])
            u = rest_u(self.theta[s], self.theta_run[s], self.tables[s], (self.step, (1 << 32) | r, 0, 0), key)
            sp.append(s)
            ids.append(self.next_id)
            self.next_id += 1
            xs.append(tuple(pos))
            ps.append(boost(self.flow, self.mass[s], u))
            kind.append(1)
        self.step += 1
        return {"species": np.array(sp, np.int64), "ids": np.array(ids, np.int64),
                "x": np.array(xs, np.float64).reshape(-1, 3), "p": np.array(ps, np.float64).reshape(-1, 3),
                "kind": np.array(kind, np.int64)}
