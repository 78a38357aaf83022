"""Simulation-shaped host driver over the C ABI.

Mirrors the reference's `Simulation` call surface for the implicit hot path
(PIC++/simulation.cpp:205-390): the method names below are the reference's, and
they run in the reference's order inside `step()`. State lives on the GPU; host
copies are made only by the explicit accessors.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

from . import _lib
from .workloads import InitialState, Workload, species_arrays

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_F = ctypes.POINTER(ctypes.c_float)

PRECISIONS = {"fp64": _lib.PRECISION_FP64, "fp32": _lib.PRECISION_FP32}
STOP_RULES = {"reference": _lib.STOP_REFERENCE, "true_residual": _lib.STOP_TRUE_RESIDUAL}
ORTHOS = {"mgs": _lib.ORTHO_MGS, "cgs2": _lib.ORTHO_CGS2}
SOLVERS = {"fp64": _lib.SOLVER_FP64, "mixed": _lib.SOLVER_MIXED}


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class Simulation:
    def __init__(self, workload: Workload, precision: str = "fp64", stop_rule: str = "reference", ortho: str = "mgs",
                 capacity: Optional[int] = None, gmres_tolerance: Optional[float] = None,
                 gmres_max_iterations: int = 50, stream: Optional[int] = None, solver: str = "fp64",
                 pressure: str = "reference", scheme: str = "implicit"):
        self.lib = _lib.load()
        self.w = workload
        cfg = _lib.PicConfig()
        cfg.nx, cfg.ny, cfg.nz = workload.n
        cfg.dx = cfg.dy = cfg.dz = workload.dx
        cfg.dt = workload.dt
        m, q, wt, _ = species_arrays(workload)
        cfg.nspecies = len(m)
        for s in range(len(m)):
            cfg.mass[s], cfg.charge[s], cfg.weight[s] = m[s], q[s], wt[s]
        cfg.capacity = int(capacity if capacity is not None else workload.capacity_needed())
        cfg.precision = PRECISIONS[precision]
        cfg.stop_rule = STOP_RULES[stop_rule]
        cfg.ortho = ORTHOS[ortho]
        cfg.gmres_tolerance = workload.gmres_tolerance if gmres_tolerance is None else gmres_tolerance
        cfg.gmres_max_iterations = gmres_max_iterations
        cfg.solver_precision = SOLVERS[solver]
        cfg.boundary_x = _lib.BOUNDARY_WALL if workload.boundary == "wall" else _lib.BOUNDARY_PERIODIC
        cfg.pressure = {"reference": _lib.PRESSURE_REFERENCE, "cell": _lib.PRESSURE_CELL}[pressure]
        cfg.smoothing_count = workload.smoothing_count
        cfg.smoothing_parameter = workload.smoothing_parameter
        cfg.scheme = {"implicit": _lib.SCHEME_IMPLICIT, "explicit": _lib.SCHEME_EXPLICIT}[scheme]
        cfg.divergence_cleanup = int(workload.divergence_cleanup)
        cfg.filter_cut_wave_number = workload.filter_cut_wave_number
        cfg.run_kT = workload.run_kT
        for a in range(3):
            cfg.e_right[a] = workload.e_right[a]
            cfg.b_right[a] = workload.b_right[a]
        if workload.inject:
            cfg.inject = 1
            for a in range(3):
                cfg.flow[a] = workload.flow[a]
            for s, sp in enumerate(workload.species):
                cfg.ppc[s] = sp.ppc
                cfg.temperature[s] = sp.temperature
            cfg.seed = workload.seed
        self.cfg = cfg
        h = ctypes.c_void_p()
        _lib.check(self.lib.pic_create(ctypes.byref(cfg), ctypes.c_void_p(stream) if stream else None, ctypes.byref(h)),
                   None, "pic_create")
        self.h = h
        self.N = workload.N
        self.last_solve: Dict[str, float] = {}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        return _lib.check(rc, self.h, what)

    # ------------------------------------------------------------ state transfer
    def load_state(self, st: InitialState):
        E = np.ascontiguousarray(np.moveaxis(st.E, -1, 0), np.float64)
        B = np.ascontiguousarray(np.moveaxis(st.B, -1, 0), np.float64)
        self._chk(self.lib.pic_upload_fields(self.h, _dp(E), _dp(B)), "pic_upload_fields")
        sp = np.ascontiguousarray(st.species, np.int32)
        ids = np.ascontiguousarray(st.ids, np.int32)
        x = np.ascontiguousarray(st.x.T, np.float64)
        p = np.ascontiguousarray(st.p.T, np.float64)
        self._chk(self.lib.pic_upload_particles(self.h, len(sp), _ip(sp), _ip(ids), _dp(x), _dp(p)),
                  "pic_upload_particles")

    def num_particles(self) -> int:
        return int(self.lib.pic_num_particles(self.h))

    def particles(self) -> Dict[str, np.ndarray]:
        n = self.num_particles()
        sp = np.zeros(n, np.int32)
        ids = np.zeros(n, np.int32)
        x = np.zeros((3, n))
        p = np.zeros((3, n))
        self._chk(self.lib.pic_download_particles(self.h, _ip(sp), _ip(ids), _dp(x), _dp(p)), "pic_download_particles")
        return dict(species=sp, ids=ids, x=x.T.copy(), p=p.T.copy())

    def array(self, name: str) -> np.ndarray:
        which = _lib.ARRAYS[name]
        n = self.lib.pic_download_array(self.h, which, None)
        self._chk(n, "pic_download_array")
        out = np.zeros(n)
        self._chk(self.lib.pic_download_array(self.h, which, _dp(out)), "pic_download_array")
        N = self.N
        nodes = (N[0] + 1) * (N[1] + 1) * (N[2] + 1)
        cells = N[0] * N[1] * N[2]
        own = int(np.prod(self.w.n))
        if name in ("rhs", "solution"):
            nu = (self.w.n[0] + (2 if self.w.boundary == "wall" else 0), self.w.n[1], self.w.n[2])
            return np.moveaxis(out.reshape(3, *nu), 0, -1)
        if n % nodes == 0 and name in ("E", "Etheta", "Enew", "J", "divPi", "mu"):
            c = n // nodes
            a = np.moveaxis(out.reshape(c, N[0] + 1, N[1] + 1, N[2] + 1), 0, -1)
            return a.reshape(a.shape[:3] + (3, 3)) if c == 9 else a
        c = n // cells
        a = np.moveaxis(out.reshape(c, *N), 0, -1)
        if name == "Pi":
            return a.reshape(a.shape[:3] + (3, 3))
        if c == 1:
            return a[..., 0]
        return a

    def device_array(self, name: str):
        ptr = _D()
        cnt = ctypes.c_longlong()
        self._chk(self.lib.pic_device_array(self.h, _lib.ARRAYS[name], ctypes.byref(ptr), ctypes.byref(cnt)),
                  "pic_device_array")
        return ctypes.cast(ptr, ctypes.c_void_p).value, cnt.value

    # ------------------------------------------------------------ reference phases
    def initialize(self):
        """Pre-loop sequence (PIC++/simulation.cpp:86-123)."""
        self._chk(self.lib.pic_initialize(self.h), "pic_initialize")

    def updateElectroMagneticParameters(self):
        """evaluateParticlesRotationTensor + updateElectroMagneticParameters
        (PIC++/mover.cpp:810-841, PIC++/momentum_collecting.cpp:365-758), then the workload's
        smoothing passes (PIC++/simulation.cpp:213-217)."""
        self._chk(self.lib.pic_rotation_and_moments(self.h), "pic_rotation_and_moments")

    evaluateParticlesRotationTensor = updateElectroMagneticParameters

    def evaluateMaxwellEquationMatrix(self):
        """Right part of the Maxwell system (PIC++/fields.cpp:441-600); the operator is matrix-free."""
        self._chk(self.lib.pic_field_rhs(self.h), "pic_field_rhs")

    def evaluateElectricField(self) -> Dict[str, float]:
        """GMRES solve for E^{n+theta} and newE (PIC++/fields.cpp:152-376)."""
        info = _lib.PicSolveInfo()
        self._chk(self.lib.pic_maxwell_solve(self.h, ctypes.byref(info)), "pic_maxwell_solve")
        self.last_solve = _info_dict(info)
        return self.last_solve

    def evaluateMagneticField(self):
        """newB = B - c' dt curl E^{n+theta} (PIC++/fields.cpp:2717-2773)."""
        self._chk(self.lib.pic_update_b(self.h), "pic_update_b")

    def moveParticles(self):
        """Mover + periodic exchange + updateDensityParameters (PIC++/mover.cpp:21-55, :1002-1210,
        PIC++/momentum_collecting.cpp:760-1020)."""
        self._chk(self.lib.pic_push(self.h), "pic_push")

    def updateFields(self):
        """E <- newE, B <- newB (PIC++/fields.cpp:422-439)."""
        self._chk(self.lib.pic_commit_fields(self.h), "pic_commit_fields")

    def updateEnergy(self) -> np.ndarray:
        """[particle, E, B, total energy, momentum x, y, z] (PIC++/simulation.cpp:1506-1640)."""
        out = np.zeros(7)
        self._chk(self.lib.pic_energy(self.h, _dp(out)), "pic_energy")
        return out

    def step(self) -> Dict[str, float]:
        info = _lib.PicSolveInfo()
        self._chk(self.lib.pic_step(self.h, ctypes.byref(info)), "pic_step")
        self.last_solve = _info_dict(info)
        return self.last_solve

    # ------------------------------------------------------------ explicit (BUNEMAN) scheme
    def load_yee(self, fields: Dict[str, np.ndarray]):
        """Yee arrays Ex Ey Ez Bx By Bz (workloads.yee_dims layouts), then the reference's
        pre-loop exchanges (PIC++/simulation.cpp:83-85)."""
        from .workloads import YEE_NAMES, yee_dims
        keep = []
        for name in YEE_NAMES:
            a = np.ascontiguousarray(fields[name], np.float64)
            if a.shape != yee_dims(self.w, name):
                raise ValueError("%s: shape %s, expected %s" % (name, a.shape, yee_dims(self.w, name)))
            keep.append(a)
        ptrs = (_lib._D * 6)(*[a.ctypes.data_as(_lib._D) for a in keep])
        self._chk(self.lib.pic_upload_yee(self.h, ptrs), "pic_upload_yee")

    def yee(self, name: str) -> np.ndarray:
        """A Yee array (Ex .. Bz, Jx .. Jz) in its ghosted layout."""
        from .workloads import yee_dims
        which = _lib.ARRAYS[name]
        n = self.lib.pic_download_array(self.h, which, None)
        self._chk(n, "pic_download_array")
        out = np.zeros(n)
        self._chk(self.lib.pic_download_array(self.h, which, _dp(out)), "pic_download_array")
        return out.reshape(yee_dims(self.w, name))

    def explicit_step(self):
        """One step of the reference's BUNEMAN loop body (PIC++/simulation.cpp:163-197)."""
        self._chk(self.lib.pic_explicit_step(self.h), "pic_explicit_step")

    def save_checkpoint(self, path: str):
        """Lossless checkpoint of the state at a step boundary (the library's own format)."""
        self._chk(self.lib.pic_checkpoint_save(self.h, os.fsencode(path)), "pic_checkpoint_save")

    def load_checkpoint(self, path: str):
        """Restores a checkpoint written by a context of the same configuration."""
        self._chk(self.lib.pic_checkpoint_load(self.h, os.fsencode(path)), "pic_checkpoint_load")

    @staticmethod
    def _backup_info(info) -> "_lib.PicBackupInfo":
        bi = _lib.PicBackupInfo()
        for key, val in info.items():
            cur = getattr(bi, key)
            if isinstance(cur, ctypes.Array):
                for i, v in enumerate(val):
                    cur[i] = v
            else:
                setattr(bi, key, val)
        return bi

    def general(self, plasma_period: float = 1.0, scale_factor: float = 1.0, **info) -> np.ndarray:
        """The numbers of a general.dat row (outputGeneral, PIC++/output.cpp:3215-3250) for the
        current state, in the row's order without the iteration and the time (PIC_GENERAL_VALUES
        of them, the order of include/picpp_amd.h). Keyword arguments as outputSimulationBackup's
        (B0 is used)."""
        bi = self._backup_info(dict(info, plasma_period=plasma_period, scale_factor=scale_factor))
        out = np.zeros(_lib.GENERAL_VALUES)
        self._chk(self.lib.pic_general(self.h, ctypes.byref(bi), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                  "pic_general")
        return out

    def outputGeneral(self, general_file: str, plasma_period: float = 1.0, scale_factor: float = 1.0, **info):
        """Appends the general.dat row of the current state in the reference's format (time and
        current_iteration from the keyword arguments)."""
        bi = self._backup_info(dict(info, plasma_period=plasma_period, scale_factor=scale_factor))
        self._chk(self.lib.pic_write_general(self.h, ctypes.byref(bi), os.fsencode(general_file)), "pic_write_general")

    def outputSimulationBackup(self, general_file: str, e_file: str, b_file: str, particles_file: str, **info):
        """The reference's backup files (PIC++/output.cpp:3432-3602) of the current state. The
        keyword arguments fill pic_backup_info (input_type, particles_per_bin, concentration,
        temperature, plasma_period, scale_factor, time, max_time, current_iteration,
        max_iteration, size, debug_mode, maxwell_matrix_size, ext_j, V0, E0, B0, species_type,
        particle_dx); missing ones are 0."""
        bi = self._backup_info(info)
        paths = [os.fsencode(p) for p in (general_file, e_file, b_file, particles_file)]
        self._chk(self.lib.pic_write_backup(self.h, ctypes.byref(bi), *paths), "pic_write_backup")

    def cleanupDivergence(self) -> Dict[str, float]:
        """cleanupDivergence(newEfield, chargeDensity) (PIC++/divergencecleanup.cpp:20-392)."""
        info = _lib.PicSolveInfo()
        self._chk(self.lib.pic_cleanup_divergence(self.h, ctypes.byref(info)), "pic_cleanup_divergence")
        return _info_dict(info)

    def cleanupDivergenceMagnetic(self) -> Dict[str, float]:
        """cleanupDivergenceMagnetic() on newB (PIC++/divergencecleanup.cpp:939-1306)."""
        info = _lib.PicSolveInfo()
        self._chk(self.lib.pic_cleanup_divergence_magnetic(self.h, ctypes.byref(info)), "pic_cleanup_divergence_magnetic")
        return _info_dict(info)

    def filterFields(self, cut_wave_number: int):
        """filterFields(cutWaveNumber) on newE and newB (PIC++/fields_filtering.cpp:21-295)."""
        self._chk(self.lib.pic_filter_fields(self.h, int(cut_wave_number)), "pic_filter_fields")

    def synchronize(self):
        self._chk(self.lib.pic_synchronize(self.h), "pic_synchronize")

    def philox(self, counter, key, n: int, out_ptr: int):
        """Philox4x64-10 outputs of counters (c0, c1 + i, c2, c3), i < n, into the device
        buffer out_ptr (4 n 64-bit words): the generator of the injection."""
        c = [int(v) for v in counter]
        k = [int(v) for v in key]
        self._chk(self.lib.pic_philox(self.h, c[0], c[1], c[2], c[3], k[0], k[1], int(n), ctypes.c_void_p(out_ptr)),
                  "pic_philox")

    # ------------------------------------------------------------ solver boundary
    def apply_operator(self, v_ptr: int, y_ptr: int):
        self._chk(self.lib.pic_apply_operator(self.h, ctypes.cast(v_ptr, _D), ctypes.cast(y_ptr, _D)),
                  "pic_apply_operator")

    def apply_operator_factored(self, v_ptr: int, y_ptr: int, precision: str = "fp64"):
        """y = A v in the factored form of the mixed-precision solver (device vectors of doubles
        for "fp64", of floats for "fp32")."""
        self._chk(self.lib.pic_apply_operator_factored(self.h, PRECISIONS[precision], ctypes.c_void_p(v_ptr),
                                                       ctypes.c_void_p(y_ptr)), "pic_apply_operator_factored")

    def gmres_solve(self, rhs_ptr: int, x_ptr: int) -> Dict[str, float]:
        info = _lib.PicSolveInfo()
        self._chk(self.lib.pic_gmres_solve(self.h, ctypes.cast(rhs_ptr, _D), ctypes.cast(x_ptr, _D), ctypes.byref(info)),
                  "pic_gmres_solve")
        return _info_dict(info)

    # ------------------------------------------------------------ learned preconditioner
    def precond_create(self):
        self._chk(self.lib.pic_precond_create(self.h), "pic_precond_create")

    def precond_params(self):
        n = int(self.lib.pic_precond_param_count())
        p = np.zeros(n, np.float32)
        ms = ctypes.c_double()
        self._chk(self.lib.pic_precond_get_params(self.h, p.ctypes.data_as(_F), ctypes.byref(ms)), "pic_precond_get_params")
        return p, ms.value

    def set_precond_params(self, params, mu_scale: float):
        p = np.ascontiguousarray(params, np.float32)
        self._chk(self.lib.pic_precond_set_params(self.h, p.ctypes.data_as(_F), float(mu_scale)), "pic_precond_set_params")

    def precond_apply(self, r_ptr: int, mu_ptr: int, z_ptr: int):
        self._chk(self.lib.pic_precond_apply(self.h, r_ptr, mu_ptr, z_ptr), "pic_precond_apply")

    def precond_gradient(self, r_ptr: int, mu_ptr: int, e_ptr: int):
        n = int(self.lib.pic_precond_param_count())
        g = np.zeros(n, np.float32)
        loss = ctypes.c_double()
        self._chk(self.lib.pic_precond_gradient(self.h, r_ptr, mu_ptr, e_ptr, ctypes.byref(loss), g.ctypes.data_as(_F)),
                  "pic_precond_gradient")
        return loss.value, g

    def precond_log(self, max_pairs: int) -> int:
        rc = self.lib.pic_precond_log(self.h, int(max_pairs))
        if max_pairs < 0:
            return int(rc)
        self._chk(rc, "pic_precond_log")
        return 0

    def precond_train(self, epochs: int, batch: int = 8, lr: float = 1e-3, holdout: float = 0.1):
        tl, hl = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.pic_precond_train(self.h, int(epochs), int(batch), float(lr), float(holdout), ctypes.byref(tl),
                                             ctypes.byref(hl)), "pic_precond_train")
        return tl.value, hl.value

    def precond_use(self, on: bool):
        self._chk(self.lib.pic_precond_use(self.h, 1 if on else 0), "pic_precond_use")


def _info_dict(info) -> Dict[str, float]:
    return dict(iterations=info.iterations, restarts=info.restarts, estimated_residual=info.estimated_residual,
                true_residual=info.true_residual, seconds=info.seconds, preconditioned=info.preconditioned,
                fallbacks=info.fallbacks)


def open_synthetic_box(w, torch, dev, **kw):
    """A context on the current torch stream holding w's synthetic initial state (E = 0, B = B0
    uniform, particles generated on the device), after the pre-loop sequence."""
    from . import workloads as wl

    stream = torch.cuda.current_stream(dev)
    sim = Simulation(w, capacity=w.capacity_needed(), stream=stream.cuda_stream, **kw)
    N = w.N
    E = np.zeros((N[0] + 1, N[1] + 1, N[2] + 1, 3))
    E[...] = np.array(w.e0)
    if w.boundary == "wall":  # no tangential E at and behind the wall (PIC++/initialization.cpp:5282-5287)
        E[:3, :, :, 1:] = 0.0
    B = np.zeros(N + (3,))
    B[...] = np.array(w.b0)
    st = wl.InitialState(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros((0, 3)), E, B)
    sim.load_state(st)
    x, p, counts = wl.make_state_on_device(w, torch, dev)
    torch.cuda.synchronize()
    cnt = (ctypes.c_longlong * len(counts))(*counts)
    _D = ctypes.POINTER(ctypes.c_double)
    rc = sim.lib.pic_upload_particles_device(sim.h, cnt, ctypes.cast(x.data_ptr(), _D), ctypes.cast(p.data_ptr(), _D), None)
    sim._chk(rc, "pic_upload_particles_device")
    del x, p
    torch.cuda.empty_cache()
    sim.initialize()
    return sim
