"""Benchmark of the MI355X bla time step (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: the Re_tau ~ 180 plane channel, 128 x 97 x 128 (nx, ny = nyp, nz),
lx = 4 pi, lz = 4 pi/3, Re_cl = 4200, constant mass flux, RK3-CN at fixed
dt = 0.005, CFL evaluated every full step as bla.i icfl=3 does.  The initial
field is the seeded divergence-free perturbed Poiseuille flow (10% rms over
|kx|,|kz| <= 16, seed 2); the state is resident on the GPU before timing
starts.  A step is one full RK3 step (three substeps, each step 2 over all
y planes and step 3 over all z planes).

Roofline: per-kernel durations come from HIP events recorded on the
library's stream around every launch of the timed region; the dominant
kernel family is reported with its algorithmic bytes (HBM-bound kernels) or
flops (the Chebyshev transform products) per launch.

CPU baseline: the reference Simson bla built from its Fortran sources
(oracle/_ref, OpenMP, 16 threads) on the same seeded field, timed by bla's
own "Average time step" over a bounded sample of steps on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "channel time-steps/s @128x97x128: FP32-mixed vs FP64 GPU at ≤1e-5 rel err"
NX, NY, NZ = 128, 97, 128
RE, XL, ZL, DT = 4200.0, 4 * math.pi, 4 * math.pi / 3, 0.005
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
REF_BIN = os.path.join(HERE, "oracle", "_ref", "bla_128x97x128_omp16")


def alg_per_launch(mixed: bool):
    """Algorithmic HBM bytes (or flops) of one launch of each kernel family.

    F = nz*ny*nx/2 complex values per field, Fp = ny*nzp*nx/2 per padded
    z-physical field; a stored complex value is 16 bytes in FP64 and 8 bytes
    in the FP32 copies of the mixed form (step 2 and the transforms), while
    the partial right-hand sides p_v, p_om stay FP64; the z-oddball plane is
    skipped by step 3 (nz-1 planes).
    """
    nxh, nzp = NX // 2, NZ + NZ // 2
    F = NZ * NY * nxh
    Fp = NY * nzp * nxh
    F3 = (NZ - 1) * NY * nxh
    b = 8 if mixed else 16
    return {
        # read u,v,w,om_x,om_z; write six z-synthesised fields
        "zsyn": ("hbm", (5 * F + 6 * Fp) * b),
        # read six fields, write three products
        "xrow": ("hbm", (6 * Fp + 3 * Fp) * b),
        # read three products, write H into three storage slots
        "zana": ("hbm", (3 * Fp + 3 * F) * b),
        # read H (3 fields), write u,v,w,om_x,om_z; read and write p_v, p_om
        # (FP64); read the per-column tau factors (gamma, 1/pivot) of lam^2 and
        # k^2 and the homogeneous solution v_h (40 B per element)
        "tau": ("hbm", (3 * F3 + 5 * F3) * b + 4 * F3 * 16 + F3 * 40),
        # dense ny x ny transform of 3 (fwd) / 5 (bwd) complex fields
        "cheb_fwd": ("flops", 2.0 * NY * NY * 2 * nxh * (NZ - 1) * 3),
        "cheb_bwd": ("flops", 2.0 * NY * NY * 2 * nxh * (NZ - 1) * 5),
    }


def traffic_file(mixed: bool) -> str:
    form = "mixed" if mixed else "fp64"
    return os.path.join(HERE, "profiles", f"hbm_traffic_step_128x97x128_{form}.md")


def profiled_traffic(family: str, mixed: bool):
    """HBM bytes per launch of a kernel family from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of that form (copy-calibrated), or None."""
    try:
        text = open(traffic_file(mixed)).read()
    except OSError:
        return None
    m = re.search(r"Machine-readable \(bytes per launch\): (.*)", text)
    if not m:
        return None
    for item in m.group(1).split(", "):
        k, _, v = item.rpartition("=")
        if k.strip() == family:
            return int(v)
    return None


def cpu_baseline(nsteps: int = 300, threads: int = 16):
    """Time the reference bla on the host (bounded sample of full steps)."""
    if not os.path.exists(REF_BIN):
        return None
    from simson_amd import blai, fields, ufile
    work = tempfile.mkdtemp(prefix="simson_cpu.")
    try:
        f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=16, rms=0.1, seed=2)
        ufile.write_u(os.path.join(work, "init.u"), f)
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1e6, maxit=3 * nsteps,
                           dt=DT, nst=3, xl=XL, cflux=True, icfl=3)
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(p))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        cmd = f"ulimit -s unlimited; exec {REF_BIN}"
        r = subprocess.run(["bash", "-c", cmd], cwd=work, env=env, capture_output=True,
                           text=True, timeout=600)
        m = re.search(r"Average time step\s*:\s*([0-9.Ee+-]+)", r.stdout)
        if r.returncode != 0 or not m:
            return None
        sec = float(m.group(1))
        return {"value": 1.0 / sec, "unit": "time-steps/s", "cores": threads,
                "kind": "reference",
                "sample": f"reference bla (Fortran, amdflang -O3 OpenMP) {NX}x{NY}x{NZ}, "
                          f"{nsteps} RK3 steps, bla's 'Average time step' = {sec:.4f} s"}
    except (subprocess.TimeoutExpired, OSError):
        return None
    finally:
        shutil.rmtree(work, ignore_errors=True)


FP32_PEAK_TFLOPS = 157.3


def time_form(precision: int, steps: int, warmup: int, model: str = ""):
    """Run one form of the step on the bench workload; returns steps/s,
    seconds, per-family (total ms, launches) and the last CFL*dt.
    model: "" (DNS), "adm-rt" (the relaxation model of sgs.i iles = 1 with
    the default sgs.i), "smag" (the dynamic Smagorinsky model of sgs.i
    iles = 2, cs = 0, ineg = 1) or "learned" (the closure network evaluated at every
    point of the dealiased grid every substep; its cost does not depend on
    the weights, so the network keeps its seeded initial weights)."""
    import numpy as np

    from simson_amd import fields
    from simson_amd.gpu import BlaGPU

    f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=16, rms=0.1, seed=2)
    g = BlaGPU(NX, NY, NZ, RE, XL, ZL, cflux=True, precision=precision)
    try:
        if model == "adm-rt":
            g.les(iles=1, cutoff=2.0 / 3.0, iord=5, chi=0.2)
        elif model == "smag":
            g.les(iles=2, cs=0.0, ineg=1)
        elif model == "learned":
            from simson_amd.sgsnet import SgsNet
            with SgsNet(1) as net:
                params = net.params
            utau2 = (180.0 / RE) ** 2
            y = np.cos(np.pi * np.arange(NY) / (NY - 1))
            g.closure(params, np.full(9, 1.0 / (utau2 * RE)), np.zeros(9),
                      (1.0 - np.abs(y)) * 180.0 / 100.0, np.full(6, 0.1 * utau2), np.zeros(6))
        g.upload_velocity(f.u, -2.0 / RE)
        g.run(warmup, DT, icfl=3, first=True)
        g.sync()
        g.timing(True)
        g.timing_read()                     # drop anything recorded so far
        t0 = time.perf_counter()
        cfl = g.run(steps, DT, icfl=3, first=False)
        g.sync()
        t1 = time.perf_counter()
        kern = g.timing_read()
        g.timing(False)
    finally:
        g.close()
    return steps / (t1 - t0), t1 - t0, kern, cfl


def time_entry_points(reps: int = 5):
    """Milliseconds per call of the FP64 statistics and pressure entry points
    on the bench workload (boxamp, boxxys 1-15, nonlinp + linearp), each
    call synchronising; not part of the step."""
    import numpy as np

    from simson_amd import fields
    from simson_amd.gpu import SIMSON_FP64, BlaGPU

    f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=16, rms=0.1, seed=2)
    g = BlaGPU(NX, NY, NZ, RE, XL, ZL, cflux=True, precision=SIMSON_FP64)
    out = {}
    try:
        g.upload_velocity(f.u, -2.0 / RE)
        g.run(1, DT, icfl=0, first=True)
        acc = np.zeros((15, NY, NX))
        for name, call in (("amp", g.amp), ("xys", lambda: g.xys(DT, acc)),
                           ("pressure", g.pressure)):
            call()                          # first call allocates
            g.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            g.sync()
            out[name] = round(1e3 * (time.perf_counter() - t0) / reps, 4)
    finally:
        g.close()
    return out


def kernel_table(kern: dict, mixed: bool):
    alg = alg_per_launch(mixed)
    flop_peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    total = max(1e-12, sum(v[0] for v in kern.values()))
    kernels = {}
    for name, (ms, cnt) in kern.items():
        if cnt == 0:
            continue
        avg_us = 1e3 * ms / cnt
        entry = {"launches": cnt, "avg_us": round(avg_us, 3), "share": round(ms / total, 4)}
        if name in alg:
            kind, amount = alg[name]
            if kind == "hbm":
                gbs = amount / (avg_us * 1e-6) / 1e9
                entry.update(bound="hbm", alg_bytes=int(amount), achieved_GBs=round(gbs, 1),
                             frac=round(gbs / HBM_PEAK_GBS, 4))
            else:
                tf = amount / (avg_us * 1e-6) / 1e12
                entry.update(bound="mfma", alg_flops=int(amount), achieved_TFLOPs=round(tf, 3),
                             frac=round(tf / flop_peak, 4))
        kernels[name] = entry
    return kernels


def roofline_of(kernels: dict, kern: dict, mixed: bool):
    alg = alg_per_launch(mixed)
    dom = max((k for k in kernels if k in alg), key=lambda k: kern[k][0])
    d = kernels[dom]
    if d["bound"] == "hbm":
        return {"kernel": dom, "bound": "hbm", "achieved": d["achieved_GBs"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"],
                "traffic": profiled_traffic(dom, mixed)}
    return {"kernel": dom, "bound": "mfma", "achieved": d["achieved_TFLOPs"],
            "peak": FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": d["frac"], "traffic": profiled_traffic(dom, mixed)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--form", choices=["both", "mixed", "fp64"], default="both",
                    help="which forms of the step to time (the headline value is the "
                         "mixed form; 'fp64' reports the FP64 step as the value)")
    ap.add_argument("--models", default="adm-rt,smag,learned",
                    help="subgrid models also timed on the mixed step (comma list of "
                         "adm-rt, smag, learned; empty for none), reported under forms")
    ap.add_argument("--no-entry-points", dest="entry_points", action="store_false",
                    help="skip timing the statistics and pressure entry points")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy a device buffer of this size three times "
                         "(a plain copy of known bytes that scales rocprofv3 --pmc "
                         "FETCH_SIZE/WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs the step on one GPU; --gpus {args.gpus} is not supported "
              "(the step is not sharded across GPUs).", file=sys.stderr)
        sys.exit(2)

    # the CPU leg runs first, before this process touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_steps)

    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED

    res = {}
    if args.form in ("both", "fp64"):
        res["fp64"] = time_form(SIMSON_FP64, args.steps, args.warmup)
    if args.form in ("both", "mixed"):
        res["mixed"] = time_form(SIMSON_MIXED, args.steps, args.warmup)
        for m in [m for m in args.models.split(",") if m]:
            res["mixed+" + m] = time_form(SIMSON_MIXED, max(1, args.steps // 2), args.warmup, m)
    entry = time_entry_points() if args.entry_points else None
    if args.calib_copy_mib > 0:
        import ctypes
        from simson_amd import build as _b
        cal = ctypes.CDLL(_b.CALIB_LIB)
        cal.simson_calib_copy.restype = ctypes.c_int
        cal.simson_calib_copy.argtypes = [ctypes.c_size_t, ctypes.c_int]
        if cal.simson_calib_copy(args.calib_copy_mib << 20, 3) != 0:
            print("calibration copy failed", file=sys.stderr)

    head = "mixed" if "mixed" in res else "fp64"
    mixed = head == "mixed"
    sps, elapsed, kern, cfl = res[head]
    kernels = kernel_table(kern, mixed)
    forms = {}
    for name, (fs, fel, fk, fc) in res.items():
        forms[name] = {"steps_per_s": round(fs, 3), "ms_per_step": round(1e3 / fs, 4),
                       "grid_points_steps_per_s": round(fs * NX * NY * NZ, 1),
                       "cfl_dt": fc, "kernels": kernel_table(fk, name.startswith("mixed"))}
    out = {
        "metric": METRIC,
        "value": round(sps, 3),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "Re_tau~180 plane channel 128x97x128, RK3-CN dt=0.005, "
                               "constant flux, " + ("FP32-mixed step" if mixed else "FP64 step"),
                   "nx": NX, "ny": NY, "nz": NZ, "re_cl": RE, "lx": "4pi", "lz": "4pi/3",
                   "dt": DT, "icfl": 3, "form": head, "parallelism": "single GPU"},
        "grid_points_steps_per_s": round(sps * NX * NY * NZ, 1),
        "mixed_vs_fp64": (round(res["mixed"][0] / res["fp64"][0], 3)
                          if "mixed" in res and "fp64" in res else None),
        "cfl_dt": cfl,
        "roofline": roofline_of(kernels, kern, mixed),
        "forms": forms,
        "entry_points_ms": entry,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
