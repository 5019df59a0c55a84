"""The C-ABI library: it loads without a GPU, exports every function
include/wf_abi.h declares, and its host-only entry points work here."""
import ctypes as C
import os
import re

import numpy as np

from weldformgpu_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "wf_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wf_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _abi.load()
    names = header_functions()
    assert len(names) >= 40
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(_abi.declared_symbols()) == names


def test_struct_sizes_match_header_layout():
    # wf_config and wf_stats are passed by pointer: their sizes must agree
    # with the C compiler's layout of include/wf_abi.h
    import subprocess
    import tempfile
    src = '#include "wf_abi.h"\n#include <stdio.h>\nint main(){printf("%zu %zu %zu\\n",' \
          'sizeof(wf_config), sizeof(wf_stats), sizeof(wf_material_params));return 0;}\n'
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == C.sizeof(_abi.wf_config)
    assert int(out[1]) == C.sizeof(_abi.wf_stats)
    assert int(out[2]) == C.sizeof(_abi.wf_material_params)


def test_field_shapes():
    lib = _abi.load()
    for name, (fid, comps, dt) in _abi.FIELDS.items():
        c, e = C.c_int(), C.c_int()
        assert lib.wf_field_shape(fid, C.byref(c), C.byref(e)) == 0
        assert c.value == comps and e.value == np.dtype(dt).itemsize, name
    assert lib.wf_field_shape(999, None, None) == _abi.WF_E_ARG


def test_config_default_matches_python_twin():
    lib = _abi.load()
    c = _abi.wf_config()
    lib.wf_config_default(C.byref(c))
    p = _abi.default_config()
    for k, _ in _abi.wf_config._fields_:
        if k == "mat":
            continue
        assert getattr(c, k) == getattr(p, k), k


def test_create_rejects_bad_config():
    lib = _abi.load()
    c = _abi.default_config()
    h = C.c_void_p()
    c.n_particles = 0
    assert lib.wf_create(C.byref(c), C.byref(h)) == _abi.WF_E_ARG
    c.n_particles = 10
    c.abi_version = 99
    assert lib.wf_create(C.byref(c), C.byref(h)) == _abi.WF_E_ARG
    assert b"abi_version" in lib.wf_last_error()
