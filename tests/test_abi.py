"""The C ABI library loads and exports every entry point include/picpp_amd.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from picpp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "picpp_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pic_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_phase_entry_points():
    syms = declared_symbols()
    for name in ("pic_create", "pic_destroy", "pic_rotation_and_moments", "pic_maxwell_solve", "pic_update_b",
                 "pic_push", "pic_commit_fields", "pic_energy", "pic_step", "pic_gmres_solve"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_bad_config_is_rejected_without_exiting():
    lib = _lib.load()
    cfg = _lib.PicConfig()
    cfg.nx = cfg.ny = cfg.nz = 3  # below the 5-cell minimum
    h = ctypes.c_void_p()
    assert lib.pic_create(ctypes.byref(cfg), None, ctypes.byref(h)) == -1
    assert not h.value
