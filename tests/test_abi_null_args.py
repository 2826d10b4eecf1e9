"""Every C-ABI entry point called with null pointers and zero sizes returns a
status and never crashes the process: on the CPU (device calls on a machine
without a GPU come back as errors) and with a device, where the calls get
past device set-up and only the host-side checks stand between them and the
kernels."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_CHILD = r"""
import ctypes, json, sys
sys.path.insert(0, sys.argv[1])
import provekit_amd as pk
L = ctypes.CDLL(pk.LIB_PATH)
names = sorted(n for n in pk.SIGNATURES)
for name in names:
    argtypes = pk.SIGNATURES[name]
    f = getattr(L, name)
    f.argtypes = argtypes
    f.restype = ctypes.c_int
    args = []
    for t in argtypes:
        if t in (ctypes.c_double, ctypes.c_float):
            args.append(0.0)
        elif t in (ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint8):
            args.append(0)
        else:
            args.append(None)  # every pointer type: NULL
    print(json.dumps({"calling": name}), flush=True)
    rc = f(*args)
    print(json.dumps({"name": name, "rc": rc}), flush=True)
print(json.dumps({"done": len(names), "device": int(L.pkgpu_device_synchronize() == 0)}), flush=True)
"""


def _run_child():
    import provekit_amd as pk
    assert os.path.exists(pk.LIB_PATH), "build the library first"
    r = subprocess.run([sys.executable, "-c", _CHILD, ROOT], capture_output=True, text=True,
                       timeout=600)
    lines = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    last = lines[-1] if lines else {}
    assert r.returncode == 0, f"crashed in {last.get('calling')}: rc={r.returncode}\n{r.stderr[-2000:]}"
    assert "done" in last and last["done"] == len(pk.SIGNATURES)
    rcs = {d["name"]: d["rc"] for d in lines if "rc" in d}
    # status codes only: 0 (nothing to do) or one of the library's error codes
    bad = {n: c for n, c in rcs.items() if not (0 <= c <= pk.ERR_INTERNAL)}
    assert not bad, bad
    return rcs, last["device"]


def test_every_entry_point_survives_null_arguments():
    _run_child()


@pytest.mark.gpu
def test_every_entry_point_survives_null_arguments_with_a_device():
    rcs, device = _run_child()
    assert device == 1, "the child process saw no GPU"
    # with a device, the zero-size device paths have nothing to do
    for name in ("pkgpu_scalar_mul_many_device", "pkgpu_sparse_mul_device", "pkgpu_stream_synchronize"):
        assert rcs[name] == 0, (name, rcs[name])
