"""The C-ABI library loads without a GPU and exports every entry point the header declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'qgate_hip.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(qghip_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert 'qghip_qubit_processor_apply_gate' in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from qgate_amd import build
    path = build.lib_path()
    if not os.path.exists(path):
        build.build()
    lib = ctypes.CDLL(path)          # no HIP call is made: loading needs no GPU
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_runtime_module_binds_every_symbol():
    """hipruntime declares ctypes signatures for every header function it uses."""
    from qgate_amd.simulator import hipruntime
    lib = hipruntime.lib()
    for name in declared_functions():
        fn = getattr(lib, name)
        assert fn.restype is not None or name == 'qghip_last_error'


def test_argument_checks_without_gpu():
    """Entry points that only check and record arguments make no HIP call."""
    import ctypes
    from qgate_amd.simulator import hipruntime
    lib = hipruntime.lib()
    h = ctypes.c_void_p()
    assert lib.qghip_qubit_processor_new(1, ctypes.byref(h)) == 0          # FP32
    p = h.value
    try:
        assert lib.qghip_qubit_processor_set_fusion(p, 1, 0, 0) == -2       # tile, low_bits 0
        assert b'low_bits' in lib.qghip_last_error()
        assert lib.qghip_qubit_processor_set_fusion(p, 1, 0, 1) == 0
        assert lib.qghip_qubit_processor_set_fusion(p, 2, 7, -1) == -2      # dense width 7
        assert lib.qghip_qubit_processor_set_fusion(p, 9, 0, -1) == -2
        assert lib.qghip_qubit_processor_set_fusion(p, 0, 0, -1) == 0
    finally:
        assert lib.qghip_qubit_processor_delete(p) == 0
    assert lib.qghip_qubit_processor_new(0, ctypes.byref(h)) == 0          # FP64: low_bits 0 is valid
    try:
        assert lib.qghip_qubit_processor_set_fusion(h.value, 1, 0, 0) == 0
    finally:
        assert lib.qghip_qubit_processor_delete(h.value) == 0
    assert lib.qghip_set_sampling_pool_workspace(3) == -2
    assert lib.qghip_set_sampling_pool_workspace(0) == 0
