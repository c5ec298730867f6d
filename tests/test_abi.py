"""The C-ABI library loads without a GPU, exports every entry point include/dpmd.h
declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, 'include', 'dpmd.h')).read()
    return sorted(set(re.findall(r'\b(dpmd_[a-z_0-9]+)\s*\(', src)))


def test_header_symbols_exported():
    from deepmd_jax import _lib
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.exported_symbols()) == set(names)


def test_version_and_error_string():
    from deepmd_jax import _lib
    L = _lib.lib()
    assert L.dpmd_version() >= 1
    assert isinstance(L.dpmd_last_error(), bytes)


def test_model_desc_validation_without_device():
    from deepmd_jax import _lib
    L = _lib.lib()
    d = _lib.ModelDesc()
    d.ntypes = 0
    h = ctypes.c_void_p()
    assert L.dpmd_model_create(ctypes.byref(d), 0, ctypes.byref(h)) == -1
    assert b'ntypes' in L.dpmd_last_error()
    d.ntypes = 2
    for i, w in enumerate((25, 60, 100)):
        d.embed_widths[i] = w
    assert L.dpmd_model_create(ctypes.byref(d), 0, ctypes.byref(h)) == -4  # 25 -> 60 not a doubling


def test_null_handles_are_rejected():
    from deepmd_jax import _lib
    L = _lib.lib()
    assert L.dpmd_compute(None, 0, None, None, None, None) == -1
    assert L.dpmd_md_run(None, None, 1, None, None, None) == -1
    assert L.dpmd_system_set_box(None, None) == -1


def test_struct_layout_matches_header(tmp_path):
    """ctypes structs have the offsets gcc gives the header's structs."""
    import subprocess
    from deepmd_jax import _lib
    src = tmp_path / 'off.c'
    fields_m = [f[0] for f in _lib.ModelDesc._fields_]
    fields_p = [f[0] for f in _lib.MDParams._fields_]
    body = ''.join('printf("%%zu\\n", offsetof(dpmd_model_desc, %s));' % f for f in fields_m)
    body += ''.join('printf("%%zu\\n", offsetof(dpmd_md_params, %s));' % f for f in fields_p)
    body += 'printf("%zu\\n%zu\\n", sizeof(dpmd_model_desc), sizeof(dpmd_md_params));'
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "dpmd.h"\nint main(){%s return 0;}\n' % body)
    exe = tmp_path / 'off'
    subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    expect = [getattr(_lib.ModelDesc, f).offset for f in fields_m] + [getattr(_lib.MDParams, f).offset for f in fields_p]
    expect += [ctypes.sizeof(_lib.ModelDesc), ctypes.sizeof(_lib.MDParams)]
    assert got == expect
