"""The learned closure on the LES configuration 128x129x128 (lx 2 pi, lz pi,
Re_cl 12000: Re_tau ~ 480 in these runs, the C4 target being ~550), a
posteriori, with the dynamic Smagorinsky model and ADM-RT as the judges.

Training data: a resolved channel at 192x129x160 in the same box (Re_cl
3300, Re_tau ~ 180, constant flux, dt 0.004) from the seeded 10%
perturbation, FP64; snapshots after 5000, 5500 and 6000 steps, filtered
with the 4-point top-hat in x and z (filter widths of about 24 and 14 wall
units, close to the LES grid's spacing at Re_tau 550: 27 and 13.5), about
7.4e5 samples in wall units (simson_amd.sgsdata.les_samples).

LES: 128x129x128 at Re_cl 12000 (FP64 step, dt 0.002, constant flux), spun
up for 4000 steps with the dynamic Smagorinsky model from the seeded
perturbation, then continued for 3000 steps with each closure (dynamic
Smagorinsky, ADM-RT, the learned network, none), the profiles averaged
over the last 2000 steps every 100 steps.  Reported: Re_tau, and the
largest deviation over y of the mean U+, u/v/w rms+ and uv+ from the
dynamic Smagorinsky run, in wall units of each run.

Asserted: every run stays finite; the learned run's Re_tau is within 10%
of the dynamic Smagorinsky run's, and none of its five profiles deviates
from that run's by more than 20%.  Measured (seeded): DNS
Re_tau 165.5; LES Re_tau 475 (dynamic Smagorinsky), 491 (ADM-RT, learned,
none).  Deviations from dynamic Smagorinsky, U / u, v, w rms / uv: ADM-RT
0.031 / 0.042 0.060 0.030 / 0.117; learned 0.030 / 0.091 0.114 0.083 /
0.143; no model 0.033 / 0.049 0.067 0.035 / 0.128.  So a posteriori at
this resolution the network trained at Re_tau 165 moves the fluctuation
profiles about twice as far from the dynamic model as ADM-RT does, while
the mean profile and Re_tau agree as well as ADM-RT's; the test records
that rather than claiming the network beats the classical closures.
"""
import math
import sys
import time

import numpy as np
import pytest

from simson_amd import fields, sgsdata
from simson_amd.profiles import profiles, wall_shear

pytestmark = pytest.mark.gpu

LX, LZ = 2 * math.pi, math.pi
DNS = dict(nx=192, ny=129, nz=160, re=3300.0, dt=0.004)
LES = dict(nx=128, ny=129, nz=128, re=12000.0, dt=0.002)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def wall_profiles(u, re):
    """Mean U+, rms+ (u, v, w) and uv+ in the wall units of the field."""
    mean, rms, uv = profiles(u)
    tw = wall_shear(mean[0], re)
    ut = math.sqrt(tw)
    return {"U": mean[0] / ut, "urms": rms[0] / ut, "vrms": rms[1] / ut,
            "wrms": rms[2] / ut, "uv": uv / (ut * ut)}, ut * re


def dev(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_learned_closure_les_550():
    from simson_amd.gpu import BlaGPU
    from simson_amd.sgsnet import closure_maps, train_closure
    d = DNS
    say(f"[gpu] C4: resolved channel {d['nx']}x{d['ny']}x{d['nz']} Re_cl {d['re']:g} for "
        "training data")
    f = fields.channel_field(d["nx"], d["ny"], d["nz"], LX, LZ, d["re"], kmax=16, rms=0.1,
                             seed=4)
    snaps = []
    with BlaGPU(d["nx"], d["ny"], d["nz"], d["re"], LX, LZ, cflux=True) as g:
        g.upload_velocity(f.u, -2.0 / d["re"])
        done = 0
        for n in (5000, 5500, 6000):
            g.run(n - done, d["dt"], icfl=3, first=done == 0)
            done = n
            snaps.append(g.velocities())
    data = [sgsdata.les_samples(s, d["nx"], d["ny"], d["nz"], LX, LZ, d["re"], fw=4)
            for s in snaps]
    X = np.concatenate([a[0] for a in data])
    T = np.concatenate([a[1] for a in data])
    re_tau_dns = float(np.mean([a[2] for a in data]))
    say(f"    DNS Re_tau {re_tau_dns:.1f}; {X.shape[0]} samples")
    assert 120.0 < re_tau_dns < 260.0, re_tau_dns
    net, sc, losses = train_closure(X, T, steps=3000, batch=4096, lr=2e-3)
    params = net.params
    net.close()
    say(f"    trained: loss {losses[0]:.3f} -> {losses[-1]:.3f}")

    e = LES
    say(f"[gpu] C4: LES {e['nx']}x{e['ny']}x{e['nz']} Re_cl {e['re']:g}: 4000 steps of "
        "dynamic Smagorinsky, then 3000 per closure")
    f = fields.channel_field(e["nx"], e["ny"], e["nz"], LX, LZ, e["re"], kmax=16, rms=0.1,
                             seed=6)
    with BlaGPU(e["nx"], e["ny"], e["nz"], e["re"], LX, LZ, cflux=True) as g:
        g.les(iles=2, cs=0.0, ineg=1)
        g.upload_velocity(f.u, -2.0 / e["re"])
        g.run(4000, e["dt"], icfl=3)
        ur, ui, _, px0 = g.download(3)
        start = ur + 1j * ui
    _, re_tau0 = wall_profiles(start, e["re"])
    say(f"    after spin-up: Re_tau {re_tau0:.1f}")
    res = {}
    for model in ("smag", "adm-rt", "learned", "none"):
        acc = None
        with BlaGPU(e["nx"], e["ny"], e["nz"], e["re"], LX, LZ, cflux=True) as g:
            if model == "smag":
                g.les(iles=2, cs=0.0, ineg=1)
            elif model == "adm-rt":
                g.les(iles=1, cutoff=2.0 / 3.0, iord=5, chi=0.2)
            elif model == "learned":
                g.closure(params, *closure_maps(sc, e["re"], re_tau0, e["ny"]))
            g.upload_velocity(start, px0)
            g.run(1000, e["dt"], icfl=3)
            g.sync()
            t0 = time.perf_counter()
            n = 0
            for _ in range(20):
                g.run(100, e["dt"], icfl=3, first=False)
                u = g.velocities()
                assert np.all(np.isfinite(u)), model
                p, rt = wall_profiles(u, e["re"])
                p["re_tau"] = rt
                acc = p if acc is None else {k: acc[k] + p[k] for k in acc}
                n += 1
            ms = 1e3 * (time.perf_counter() - t0) / 2000
        res[model] = {k: v / n for k, v in acc.items()}
        say(f"    {model:8s}: Re_tau {res[model]['re_tau']:.1f}, {ms:.2f} ms/step "
            "(with the profile downloads)")
    ref = res["smag"]
    for model in ("adm-rt", "learned", "none"):
        devs = {k: dev(res[model][k], ref[k]) for k in ("U", "urms", "vrms", "wrms", "uv")}
        res[model]["dev"] = devs
        say(f"    {model:8s} vs dynamic Smagorinsky: " +
            ", ".join(f"{k} {v:.3f}" for k, v in devs.items()))
    for model in ("smag", "adm-rt", "learned", "none"):
        r = res[model]
        say(f"    {model:8s} peaks: u_rms+ {r['urms'].max():.3f}, v_rms+ {r['vrms'].max():.3f}, "
            f"w_rms+ {r['wrms'].max():.3f}, -uv+ {-r['uv'].min():.3f}, "
            f"U+ centre {r['U'][e['ny'] // 2]:.2f}")
    assert abs(res["learned"]["re_tau"] / ref["re_tau"] - 1.0) < 0.1
    for k in ("U", "urms", "vrms", "wrms", "uv"):
        assert res["learned"]["dev"][k] <= 0.2, (k, res["learned"]["dev"])
