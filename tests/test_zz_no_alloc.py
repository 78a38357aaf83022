"""pic_energy_noalloc: updateEnergy of caller-owned device arrays with no context and no
allocation (its partial sums live in the library's static device array), against the context's
own pic_energy on the same state. Two calls in a row, the second on a smaller grid with fewer
particles, so that the second call runs over what the first left in the shared array. Both use
the same kernels' loops and block reductions in the same particle order, so the bar is the FP64
parity bar, 1e-14 normwise (measured: equal)."""
import ctypes

import numpy as np
import pytest

from conftest import relerr
from picpp_amd import _lib
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


def energy_without_context(sim, torch):
    """pic_energy_noalloc on the context's E, B and particles (copied to caller-owned buffers)."""
    sim.synchronize()
    e_ptr, _ = sim.device_array("E")
    b_ptr, _ = sim.device_array("B")
    P = sim.particles()
    sp = P["species"]
    assert np.all(np.diff(sp) >= 0)  # the context keeps its particles grouped by species
    counts = (ctypes.c_longlong * 8)(*[int(np.sum(sp == s)) for s in range(8)])
    p = torch.from_numpy(np.ascontiguousarray(P["p"].T)).cuda()
    out = torch.full((7,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = sim.lib.pic_energy_noalloc(ctypes.byref(sim.cfg), ctypes.c_void_p(e_ptr), ctypes.c_void_p(b_ptr),
                                    ctypes.c_void_p(p.data_ptr()), counts, ctypes.c_void_p(out.data_ptr()), None)
    assert rc == 0, sim.lib.pic_last_error(None)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_energy_matches_the_context_twice_second_smaller():
    import torch
    assert torch.cuda.is_available()
    results = []
    for n, ppc, steps in (((12, 10, 8), 8, 2), ((6, 5, 5), 4, 1)):
        w = wl.mini(n=n, ppc=ppc, seed=13)
        st = wl.make_initial_state(w, perturb=0.05)
        print("case pic_energy_noalloc", n, "ppc", ppc, flush=True)
        sim = Simulation(w)
        sim.load_state(st)
        sim.initialize()
        for _ in range(steps):
            sim.step()
        want = sim.updateEnergy()
        got = energy_without_context(sim, torch)
        sim.close()
        err = relerr(got, want)
        print("  context %s\n  no-alloc %s\n  relerr %.1e" % (want, got, err))
        assert np.all(np.isfinite(got))
        assert err <= 1e-14
        results.append(got)
    assert not np.array_equal(results[0], results[1])


def test_noalloc_energy_refuses_bad_arguments():
    lib = _lib.load()
    cfg = _lib.PicConfig()
    cfg.nx = cfg.ny = cfg.nz = 8
    cfg.dx = cfg.dy = cfg.dz = 1.0
    cfg.nspecies = 9
    counts = (ctypes.c_longlong * 8)()
    one = ctypes.c_void_p(1)
    assert lib.pic_energy_noalloc(ctypes.byref(cfg), one, one, one, counts, one, None) == _lib.ERRORS_BY_NAME["argument"]
    assert b"species" in lib.pic_last_error(None)
