"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product path never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.oracle_create.restype = ctypes.c_void_p
    lib.oracle_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.c_int, _D, _D, _D, _I]
    lib.oracle_destroy.argtypes = [ctypes.c_void_p]
    lib.oracle_set_fields.argtypes = [ctypes.c_void_p, _D, _D]
    lib.oracle_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_long, _I, _I, _D, _D]
    lib.oracle_num_particles.restype = ctypes.c_long
    lib.oracle_num_particles.argtypes = [ctypes.c_void_p]
    lib.oracle_get_particles.argtypes = [ctypes.c_void_p, _I, _I, _D, _D, _D]
    lib.oracle_get_array.restype = ctypes.c_long
    lib.oracle_get_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _D]
    lib.oracle_get_energy.argtypes = [ctypes.c_void_p, _D]
    lib.oracle_gmres_iterations.restype = ctypes.c_int
    lib.oracle_gmres_iterations.argtypes = [ctypes.c_void_p]
    for name in ("initial_sequence", "update_particle_maps", "evaluate_rotation_tensor", "update_em_parameters",
                 "build_rhs", "evaluate_electric_field", "evaluate_magnetic_field", "exchange_fields",
                 "move_particles", "update_density_parameters", "update_fields", "update_energy"):
        getattr(lib, "oracle_" + name).argtypes = [ctypes.c_void_p]
    lib.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.oracle_apply_operator.argtypes = [ctypes.c_void_p, _D, _D]
    lib.oracle_row_entries.restype = ctypes.c_int
    lib.oracle_row_entries.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4 + [_D, _I]
    lib.oracle_gmres.restype = ctypes.c_int
    lib.oracle_gmres.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_double, ctypes.c_int]
    lib.oracle_set_mu.argtypes = [ctypes.c_void_p, _D]
    lib.oracle_bspline.restype = ctypes.c_double
    lib.oracle_bspline.argtypes = [ctypes.c_double] * 4
    lib.oracle_alpha.argtypes = [ctypes.c_double, _D, ctypes.c_double, _D, _D, _D]
    for name in ("yee_init", "yee_exchange", "tristan_b_half", "tristan_move", "tristan_flux", "tristan_e",
                 "erase_and_exchange"):
        getattr(lib, "oracle_" + name).argtypes = [ctypes.c_void_p]
    lib.oracle_yee_dims.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _I]
    lib.oracle_yee_get.restype = ctypes.c_long
    lib.oracle_yee_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _D]
    lib.oracle_yee_set.restype = ctypes.c_long
    lib.oracle_yee_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _D]
    lib.oracle_explicit_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.oracle_interp_yee.argtypes = [ctypes.c_void_p, _D, _D, _D]
    _lib = lib
    return lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_I)


class OracleSim:
    """The reference's implicit step restated on the CPU in FP64 (Simulation-shaped)."""

    def __init__(self, workload, state=None):
        lib = _load()
        from picpp_amd.workloads import species_arrays

        self.w = workload
        m, q, wt, ty = species_arrays(workload)
        self._keep = (m, q, wt, ty)
        self.h = lib.oracle_create(*workload.n, workload.dx, workload.dx, workload.dx, workload.dt, len(m),
                                   _dp(m), _dp(q), _dp(wt), _ip(ty))
        if not self.h:
            raise ValueError("oracle rejected the grid (each axis needs at least 5 cells)")
        self.lib = lib
        self.N = tuple(v + 4 for v in workload.n)
        if state is not None:
            self.set_state(state)

    def __del__(self):
        if getattr(self, "h", None) and self.lib is not None:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def set_state(self, st):
        E = np.ascontiguousarray(st.E, np.float64)
        B = np.ascontiguousarray(st.B, np.float64)
        self.lib.oracle_set_fields(self.h, _dp(E), _dp(B))
        sp = np.ascontiguousarray(st.species, np.int32)
        ids = np.ascontiguousarray(st.ids, np.int32)
        x = np.ascontiguousarray(st.x, np.float64)
        p = np.ascontiguousarray(st.p, np.float64)
        self.lib.oracle_set_particles(self.h, len(sp), _ip(sp), _ip(ids), _dp(x), _dp(p))

    # ---- phases (reference method names in comments)
    def initial_sequence(self):
        self.lib.oracle_initial_sequence(self.h)

    def step(self, n=1):
        self.lib.oracle_step(self.h, n)

    def __getattr__(self, name):
        if name.startswith("_") or name in ("lib", "h"):
            raise AttributeError(name)
        fn = getattr(_load(), "oracle_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda: fn(self.h)

    # ---- state access
    def particles(self):
        n = self.lib.oracle_num_particles(self.h)
        sp = np.zeros(n, np.int32)
        ids = np.zeros(n, np.int32)
        x = np.zeros((n, 3))
        p = np.zeros((n, 3))
        a = np.zeros((n, 9))
        self.lib.oracle_get_particles(self.h, _ip(sp), _ip(ids), _dp(x), _dp(p), _dp(a))
        return dict(species=sp, ids=ids, x=x, p=p, alpha=a)

    def array(self, name: str) -> np.ndarray:
        cnt = self.lib.oracle_get_array(self.h, name.encode(), None)
        if cnt < 0:
            raise KeyError(name)
        out = np.zeros(cnt)
        self.lib.oracle_get_array(self.h, name.encode(), _dp(out))
        N = self.N
        nodes = (N[0] + 1) * (N[1] + 1) * (N[2] + 1)
        cells = N[0] * N[1] * N[2]
        if cnt % nodes == 0 and cnt // nodes in (3, 9):
            return out.reshape(N[0] + 1, N[1] + 1, N[2] + 1, *((3,) if cnt // nodes == 3 else (3, 3)))
        if cnt == cells:
            return out.reshape(N)
        if cnt == cells * 3:
            return out.reshape(N + (3,))
        if cnt == cells * 9:
            return out.reshape(N + (3, 3))
        return out

    def energy(self) -> np.ndarray:
        out = np.zeros(7)
        self.lib.oracle_get_energy(self.h, _dp(out))
        return out

    def gmres_iterations(self) -> int:
        return self.lib.oracle_gmres_iterations(self.h)

    def apply_operator(self, v: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(v, np.float64)
        out = np.zeros_like(v)
        self.lib.oracle_apply_operator(self.h, _dp(v), _dp(out))
        return out

    def gmres(self, b: np.ndarray, precision=1e-8, max_iterations=50):
        b = np.ascontiguousarray(b, np.float64).copy()
        x = np.zeros_like(b)
        it = self.lib.oracle_gmres(self.h, _dp(b), _dp(x), precision, max_iterations)
        return x, it

    def set_mu(self, mu: np.ndarray):
        mu = np.ascontiguousarray(mu, np.float64)
        self.lib.oracle_set_mu(self.h, _dp(mu))

    # ---- explicit (BUNEMAN / Tristan) path
    def yee_dims(self, name: str):
        d = np.zeros(3, np.int32)
        self.lib.oracle_yee_dims(self.h, name.encode(), _ip(d))
        return tuple(int(v) for v in d)

    def yee(self, name: str) -> np.ndarray:
        out = np.zeros(self.lib.oracle_yee_get(self.h, name.encode(), None))
        self.lib.oracle_yee_get(self.h, name.encode(), _dp(out))
        return out.reshape(self.yee_dims(name))

    def set_yee(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        assert a.shape == self.yee_dims(name), (name, a.shape)
        self.lib.oracle_yee_set(self.h, name.encode(), _dp(a))

    def explicit_step(self, n=1):
        self.lib.oracle_explicit_step(self.h, n)

    def interp_yee(self, x):
        x = np.ascontiguousarray(x, np.float64)
        E, B = np.zeros(3), np.zeros(3)
        self.lib.oracle_interp_yee(self.h, _dp(x), _dp(E), _dp(B))
        return E, B

    def row_entries(self, i, j, k, c):
        val = np.zeros(64)
        idx = np.zeros(64, np.int32)
        cnt = self.lib.oracle_row_entries(self.h, i, j, k, c, _dp(val), _ip(idx))
        return val[:cnt], idx[:cnt]


def bspline(x, hw, left, right) -> float:
    return _load().oracle_bspline(x, hw, left, right)


def alpha(beta, v, gamma, E, B) -> np.ndarray:
    v = np.ascontiguousarray(v, np.float64)
    E = np.ascontiguousarray(E, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    out = np.zeros(9)
    _load().oracle_alpha(beta, _dp(v), gamma, _dp(E), _dp(B), _dp(out))
    return out.reshape(3, 3)
