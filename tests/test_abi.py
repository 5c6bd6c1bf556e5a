"""The C-ABI library loads and exports every entry point the header declares.

CPU-only: no compute call is made (there may be no GPU here).
"""
import os
import re
import subprocess

from simson_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "simson_gpu.h")).read()
    return sorted(set(re.findall(r"\b(simson_(?:gpu|les|mlp|fringe|wall|fsc|base|parallel|local)_\w+)\s*\(", text)))


def test_header_and_python_agree():
    assert sorted(gpu.EXPORTS) == declared()


def test_library_exports_every_symbol():
    assert os.path.exists(gpu.LIBPATH), "build the library first (python -m simson_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [s for s in declared() if s not in syms]
    assert not missing, missing
    L = gpu.lib()          # loads (dlopen) without touching a device
    for s in declared():
        assert hasattr(L, s)


def test_init_rejects_unsupported_settings():
    """Argument validation happens before any device call."""
    import ctypes as C
    bad = [
        gpu.SimsonParams(32, 33, 32, 0, 1, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),   # ibc=1
        gpu.SimsonParams(32, 34, 32, 0, 0, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),   # even ny
        gpu.SimsonParams(28, 33, 32, 0, 0, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),   # nx%4
        gpu.SimsonParams(32, 33, 32, 7, 0, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),   # precision
        gpu.SimsonParams(32, 33, 28, 0, 0, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),   # nzp=42 (7)
        gpu.SimsonParams(32, 385, 32, 0, 0, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),  # ny > 383
        gpu.SimsonParams(32, 33, 32, 0, 101, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),  # 101 + cflux
        gpu.SimsonParams(32, 33, 32, 0, 100, 1, 3000.0, 12.5, 6.28, 0.0, 0.0),  # ibc=100, cflux
    ]
    # (lengths without a specialised FFT kernel are checked against the
    # device's LDS limit, so only on a device)
    for p in bad:
        h = C.c_void_p()
        assert gpu.lib().simson_gpu_init(C.byref(p), C.byref(h)) == -1
        assert not h.value
