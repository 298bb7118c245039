"""The C-ABI libraries load and export every function include/*.h declares."""
import ctypes
import os
import re

import libmesh_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmh_\w+|lmesh_\w+)\s*\(", text)))


def test_hip_library_exports_header():
    lib = ctypes.CDLL(L.LIB_HIP)
    names = declared("lmhip.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole ABI
    assert set(names) <= set(L.HIP_API), set(names) - set(L.HIP_API)


def test_host_library_exports_header():
    lib = ctypes.CDLL(L.LIB_HOST)
    names = declared("lmesh_host.h")
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(L.HOST_API)


def test_no_cpu_fallback_without_gpu():
    """Compute entry points need a device: without one, context creation fails
    loudly instead of silently computing elsewhere."""
    import subprocess
    import sys
    code = ("import libmesh_amd as L\n"
            "try:\n    L.Context(0)\n    print('ctx')\nexcept L.LmhError as e:\n    print('error', e)\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, env=env, timeout=120)
    assert "error" in r.stdout, r.stdout + r.stderr
