"""The C-ABI library loads, exports every entry point declared in
include/ether_gpu.h, and fails loudly (no CPU fallback) without a GPU."""
import ctypes
import os
import re

import pytest

from ether_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "ether_gpu.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(ether_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(E.LIB_PATH)
    names = declared()
    assert len(names) >= 38
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    bound = {n for n, _, _ in E.EXPORTS}
    assert set(declared()) == bound


def test_error_path_without_device_or_with_bad_args():
    lib = E.load_library()
    h = ctypes.c_void_p()
    rc = lib.ether_ctx_create(7, 10, 1, ctypes.c_uint64(1), ctypes.byref(h))
    assert rc == E.ETHER_ERR_ARG
    buf = ctypes.create_string_buffer(256)
    lib.ether_last_error(None, buf, 256)
    assert b"model" in buf.value


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="only meaningful on a host without a GPU")
def test_no_gpu_fails_loudly():
    with pytest.raises(E.EtherError) as ei:
        E.Engine(E.MODEL_XYZ, 8, 1, 1)
    assert ei.value.code == E.ETHER_ERR_NODEV
