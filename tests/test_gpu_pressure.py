"""The cell-local pressure tensor (pressure="cell", PIC_PRESSURE_CELL) against what the
mathematics fixes, and the reference's deposit (the default) against its known shape.

The reference adds each particle's pressure q w corr (alpha u)(alpha u)^T at
Pi[curI][j][k] with the stencil slot indices j, k in {0, 1}
(PIC++/momentum_collecting.cpp:565, :577). The cell-local form uses the particle's own
cells. Checks:
- with E = B = 0 every particle's alpha u is its velocity v; when all particles share one
  velocity the cell-local Pi must equal rho v v^T cell by cell, where rho is the charge
  density of the same cell maps (updateDensityParameters). A wrong dual-cell offset, a
  transposed component or a wrong fold breaks the proportionality on random positions;
- on a lattice of electrons whose y-z columns are identical, the cell-local Pi is the same in
  every column, while the reference's deposit puts all of it in the columns j, k in {n, n+1}
  after the folds (documenting the difference).
"""
import numpy as np
import pytest

from conftest import relerr
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


def cold_state(w, u0, lattice=False):
    st = wl.make_initial_state(w)
    if lattice:
        n = np.array(w.n)
        cells = np.stack(np.meshgrid(*[np.arange(v) for v in n], indexing="ij"), -1).reshape(-1, 3)
        xs = [n * w.dx + (np.repeat(cells, s.ppc, axis=0) + (0.37, 0.29, 0.61)) * w.dx for s in w.species]
        st.x[:] = np.concatenate(xs)
    m = np.array([w.species[s].mass for s in st.species])
    st.p[:] = m[:, None] * np.array(u0)[None, :]
    return st


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_cell_pressure_is_rho_v_v(precision, tol):
    w = wl.Workload("pcell-10x8x6", (10, 8, 6), 1.0, 0.1, 1, wl.electron_proton(4, 0.01, 100.0), seed=3)
    u0 = (0.3, -0.2, 0.5)
    v = np.array(u0) / np.sqrt(1 + np.dot(u0, u0))
    print("case", w.name, precision, flush=True)
    sim = Simulation(w, precision=precision, pressure="cell")
    sim.load_state(cold_state(w, u0))
    sim.initialize()
    rho = sim.array("density")[..., 5 * len(w.species)]
    sim.updateElectroMagneticParameters()
    Pi = sim.array("Pi")
    want = rho[..., None, None] * np.outer(v, v)[None, None, None]
    own = (slice(2, w.n[0] + 2), slice(2, w.n[1] + 2), slice(2, w.n[2] + 2))
    err = relerr(Pi[own], want[own])
    print("cell-local Pi vs rho v v^T: %.2e" % err)
    assert err <= tol
    sim.close()


def test_lattice_columns_cell_versus_reference():
    # one species: a uniform lattice deposits a uniform density, so two species would cancel
    w = wl.Workload("plat-8x8x8", (8, 8, 8), 1.0, 0.1, 1, wl.electron_proton(1, 0.01, 100.0)[:1], seed=3)
    u0 = (-0.9, 0.1, 0.0)
    print("case", w.name, flush=True)
    out = {}
    for mode in ("cell", "reference"):
        sim = Simulation(w, precision="fp64", pressure=mode)
        sim.load_state(cold_state(w, u0, lattice=True))
        sim.initialize()
        sim.updateElectroMagneticParameters()
        out[mode] = sim.array("Pi")[2:10, 2:10, 2:10]
        sim.close()
    cell = out["cell"]
    spread = np.abs(cell - cell[:, :1, :1]).max() / np.abs(cell).max()
    print("cell-local Pi: max column spread / max %.2e" % spread)
    assert spread <= 1e-13
    ref = out["reference"]
    mass = np.abs(ref).sum(axis=(0, 3, 4))  # per (j, k) column
    corner = mass[-2:, -2:].sum() / mass.sum()
    print("reference Pi: share in the columns j, k in {n, n+1}: %.6f" % corner)
    assert corner > 0.999
