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


def _c_layout(struct, fields):
    """sizeof and offsetof of a header struct, from a C program compiled against the header."""
    import subprocess
    import tempfile

    body = ['printf("%%zu\\n", sizeof(%s));' % struct]
    body += ['printf("%%zu\\n", offsetof(%s, %s));' % (struct, f) for f in fields]
    src = ('#include <stddef.h>\n#include <stdio.h>\n#include "picpp_amd.h"\nint main(void) {\n%s\nreturn 0;\n}\n'
           % "\n".join(body))
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "layout.c")
        exe = os.path.join(tmp, "layout")
        with open(c, "w") as fh:
            fh.write(src)
        subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


def test_ctypes_structs_match_the_header_layout():
    """The ctypes mirrors of pic_config, pic_solve_info and pic_backup_info have the header's
    size and every field's offset, as a C compiler lays them out."""
    for cls, name in ((_lib.PicConfig, "pic_config"), (_lib.PicSolveInfo, "pic_solve_info"),
                      (_lib.PicBackupInfo, "pic_backup_info")):
        fields = [f[0] for f in cls._fields_]
        size, offs = _c_layout(name, fields)
        assert ctypes.sizeof(cls) == size, name
        for f, off in zip(fields, offs):
            assert getattr(cls, f).offset == off, (name, f)
