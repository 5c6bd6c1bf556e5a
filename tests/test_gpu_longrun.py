"""The mixed form over 10^4 steps of the turbulent channel (128x97x128,
Re_cl 4200, lx 4 pi, lz 4 pi/3, dt 0.005, constant flux), against FP64.

The seeded field (10% rms over |kx|, |kz| <= 16) is spun up for 2000 FP64
steps.  From that state run, for 10^4 steps each: the FP64 step, the mixed
step, and two FP64 controls whose velocities are perturbed by relative
1e-7 (FP32 rounding) with different seeds (simson_amd.profiles.long_run).
Turbulence amplifies any perturbation by about 1e5 over 10^4 steps, so
instantaneous profiles of two runs that differ by rounding separate; what
stays meaningful is the statistics.  Checked:

* after 1000 steps the instantaneous mean and rms profiles of the mixed run
  are within 1e-5 of FP64;
* the profiles averaged over steps 5000-10^4 (every 250 steps) -- mean U,
  u/v/w rms, uv and the wall shear -- deviate from FP64's no more than 3x
  as far as the largest of the controls' deviations.  Once two runs have
  decorrelated, the deviation of their 50-time-unit averages is a random
  variable whose size the averaging window sets (~1e-5 for U, ~1e-4 for
  the stresses here); the controls sample it, and the mixed run is one
  more sample: it adds nothing beyond the flow's own sensitivity to a
  rounding-sized perturbation.  The runs are deterministic (bit-identical
  on repetition), so the check is too.  Measured: U 9.9e-6 (controls
  5.6e-6), u_rms 7.0e-5 (6.0e-5), v_rms 8.7e-5 (8.9e-5), w_rms 8.1e-5
  (1.1e-4), uv 1.5e-4 (1.2e-4), wall shear 1.7e-5 (7.0e-6).
"""
import math
import sys

import pytest

from simson_amd.profiles import long_run

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mixed_statistics_over_10k_steps():
    say("[gpu] 128x97x128: 2000-step FP64 spin-up, then 4 runs of 10^4 steps")
    r = long_run(128, 97, 128, 4200.0, 4 * math.pi, 4 * math.pi / 3, 0.005,
                 spinup=2000, nsteps=10000, every=250, say=say)
    say(f"    after 1000 steps, mixed vs FP64: {r['after_1000']}")
    for k, d in r["averaged"].items():
        say(f"    averaged {k}: mixed {d['mixed']:.2e}, controls {d['control']:.2e}")
    assert max(r["after_1000"].values()) < 1e-5, r["after_1000"]
    for k, d in r["averaged"].items():
        assert d["mixed"] <= 3.0 * d["control"], (k, d)
