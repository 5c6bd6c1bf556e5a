"""simson_gpu_amp_noalloc through the C ABI: the plane statistics computed
in the library's fixed scratch array equal simson_gpu_amp's (which uses the
handle's own buffer) bit for bit, on two calls in a row with the second on a
smaller grid, each against the allocating entry point on the same state."""
import math
import sys

import numpy as np
import pytest

from simson_amd import fields
from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU, lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def amp_noalloc(g):
    a = np.full((20, g.ny), np.nan)
    st = lib().simson_gpu_amp_noalloc(g.h, a)
    assert st == 0, st
    return a


@pytest.mark.parametrize("precision", [SIMSON_FP64, SIMSON_MIXED], ids=["fp64", "mixed"])
def test_amp_noalloc_equals_amp(precision):
    cases = [(64, 97, 48, 4200.0, 4 * math.pi, 4 * math.pi / 3, 10, 5e-2),
             (32, 33, 32, 3000.0, 4 * math.pi, 2 * math.pi, 8, 1e-2)]
    for nx, ny, nz, re, xl, zl, kmax, rms in cases:
        say(f"[gpu] amp_noalloc {nx}x{ny}x{nz} precision {precision}")
        f = fields.channel_field(nx, ny, nz, xl, zl, re, kmax=kmax, rms=rms, seed=3)
        with BlaGPU.from_ufield(f, precision=precision) as g:
            g.upload_velocity(f.u, -2.0 / re)
            g.run(2, 0.005, icfl=3)
            ref = g.amp()
            got = amp_noalloc(g)
            assert np.array_equal(got, ref)
            assert np.abs(ref[:8]).max() > 0


def test_amp_noalloc_rejects_large_ny():
    """ny above the scratch limit (257 points) is an argument error, checked
    before any launch; the allocating entry point serves that handle."""
    say("[gpu] amp_noalloc refuses ny = 289")
    f = fields.channel_field(16, 289, 8, 2 * math.pi, math.pi, 3000.0, kmax=2, rms=1e-2,
                             seed=3)
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / 3000.0)
        a = np.zeros((20, 289))
        assert lib().simson_gpu_amp_noalloc(g.h, a) == -1
        assert np.abs(g.amp()[:8]).max() > 0
