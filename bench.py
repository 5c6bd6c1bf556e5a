"""Benchmark of the MI355X bla time step (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Headline workload: the Re_tau ~ 180 plane channel, 128 x 97 x 128 (nx,
ny = nyp, nz), lx = 4 pi, lz = 4 pi/3, Re_cl = 4200, constant mass flux,
RK3-CN at fixed dt = 0.005, CFL evaluated every full step as bla.i icfl=3
does.  The initial field is the seeded divergence-free perturbed Poiseuille
flow (10% rms over |kx|,|kz| <= 16, seed 2); the state is resident on the GPU
before timing starts.  A step is one full RK3 step (three substeps, each
step 2 over all y planes and step 3 over all z planes).  `value` is the
mixed form's steps/s, timed with no per-kernel instrumentation; a separate
instrumented stretch of steps gives the per-kernel table.

Also measured: the FP64 step (mixed_vs_fp64), the mixed step with each
subgrid model, the large channel 384 x 257 x 384 (Re_cl 13000, lx 2 pi,
lz pi, dt 0.002, 10% rms over |kx|,|kz| <= 32) in both forms, the spatial
Blasius boundary layer 512 x 97 x 64 (Re_delta* 400 at the inflow, fringe,
free stream above, wall strip at F = 100) in both forms, and the
mixed form's error: the profile deviation from FP64 after the bench's own
steps and, over 10^4 steps from a 2000-step spun-up state, the averaged
statistics against two FP64 controls perturbed at 1e-7.

Roofline: per-kernel durations come from HIP events recorded on the
library's stream around every launch of the instrumented stretch; the
dominant kernel family is reported with its algorithmic bytes (HBM-bound)
or flops (the Chebyshev transform products) per launch, and the whole step
with the algorithmic bytes of the reference's data flow.

CPU baseline: the reference Simson bla built from its Fortran sources
(oracle/_ref, OpenMP) on the same seeded field, with one OpenMP thread per
core available to this process, timed by bla's own "Average time step" over
a bounded sample of steps on the host cores.
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
from dataclasses import dataclass

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "channel time-steps/s @128x97x128: FP32-mixed vs FP64 GPU at ≤1e-5 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# the reference bla for the 128x97x128 grid, compiled for up to 64 OpenMP
# threads (par.f nthread); run with OMP_NUM_THREADS = the cores available
REF_BIN = os.path.join(HERE, "oracle", "_ref", "bla_128x97x128_omp")
REF_NTHREAD = 64


@dataclass(frozen=True)
class Workload:
    label: str
    nx: int
    ny: int
    nz: int
    re: float
    xl: float
    zl: float
    dt: float
    kmax: int
    rms: float
    seed: int

    @property
    def tag(self):
        return f"{self.nx}x{self.ny}x{self.nz}"

    def field(self):
        from simson_amd import fields
        return fields.channel_field(self.nx, self.ny, self.nz, self.xl, self.zl, self.re,
                                    kmax=self.kmax, rms=self.rms, seed=self.seed)


CHANNEL = Workload("Re_tau~180 plane channel", 128, 97, 128, 4200.0, 4 * math.pi,
                   4 * math.pi / 3, 0.005, 16, 0.1, 2)
LARGE = Workload("Re_tau~590 plane channel", 384, 257, 384, 13000.0, 2 * math.pi, math.pi,
                 0.002, 32, 0.1, 5)


def alg_per_launch(mixed: bool, w: Workload = CHANNEL):
    """Algorithmic HBM bytes (or flops) of one launch of each kernel family.

    The bytes follow the reference's data flow per substep: step 2 reads
    u, v, w, om_x, om_z and writes H_1..3 (5 + 3 fields); step 3 reads H_1..3
    and p_v, p_om and writes u, v, w, om_x, om_z, p_v, p_om (5 + 7 fields).
    F = nz*ny*nx/2 complex values per field; a stored complex value is 16 B
    in FP64 and 8 B in the FP32 copies of the mixed form, while p_v, p_om
    stay FP64 (16 B); step 3 skips the z-oddball plane (nz-1 planes).

    * "tau": the column solver owns step 3's whole data flow (it reads H and
      p in Chebyshev space and writes the new fields' coefficients), so it
      is priced at 5 reads + 7 writes per column element.
    * "zsyn" reads step 2's 5 input fields and "zana" writes its 3 output
      fields; the z-padded planes between zsyn, xrow and zana are this
      design's intermediates and carry no algorithmic bytes ("step2" prices
      the three kernels together).
    * The Chebyshev transforms are priced in the flops of the parity-split
      products issued: per real column 2 ((H + 1)^2 + H^2) with ny = 2H + 1
      (half of the dense 2 ny^2).
    """
    nxh = w.nx // 2
    F = w.nz * w.ny * nxh
    F3 = (w.nz - 1) * w.ny * nxh
    b = 8 if mixed else 16
    H = (w.ny - 1) // 2
    split = 2.0 * ((H + 1) ** 2 + H ** 2)
    cols = 2 * nxh * (w.nz - 1)
    return {
        "zsyn": ("hbm", 5 * F * b),
        "zana": ("hbm", 3 * F * b),
        "tau": ("hbm", (3 * F3 + 5 * F3) * b + 4 * F3 * 16),
        "cheb_fwd": ("flops", split * cols * 3),
        "cheb_bwd": ("flops", split * cols * 5),
    }


def step2_bytes(mixed: bool, w: Workload = CHANNEL):
    return (5 + 3) * w.nz * w.ny * (w.nx // 2) * (8 if mixed else 16)


def step_bytes(mixed: bool, w: Workload = CHANNEL):
    """Algorithmic HBM bytes of one full RK3 step: three substeps of the
    data flow above (step 2: 8 fields; step 3: 8 fields of the stored
    precision and 4 FP64 partial right-hand-side fields)."""
    F = w.nz * w.ny * (w.nx // 2)
    b = 8 if mixed else 16
    return 3 * (8 * F * b + 8 * F * b + 4 * F * 16)


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


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nsteps: int = 300):
    """Time the reference bla on the host (bounded sample of full steps),
    one OpenMP thread per available core (at most the compiled nthread)."""
    if not os.path.exists(REF_BIN):
        return None
    from simson_amd import blai, ufile
    w = CHANNEL
    threads = min(host_cores(), REF_NTHREAD)
    work = tempfile.mkdtemp(prefix="simson_cpu.")
    try:
        ufile.write_u(os.path.join(work, "init.u"), w.field())
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1e6, maxit=3 * nsteps,
                           dt=w.dt, nst=3, xl=w.xl, cflux=True, icfl=3)
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
                "sample": f"reference bla (Fortran, amdflang -O3 OpenMP, {threads} threads on "
                          f"{cpu_model()}) {w.tag}, {nsteps} RK3 steps, bla's 'Average time "
                          f"step' = {sec:.4f} s"}
    except (subprocess.TimeoutExpired, OSError):
        return None
    finally:
        shutil.rmtree(work, ignore_errors=True)


def time_form(w: Workload, precision: int, steps: int, warmup: int, model: str = "",
              isteps: int = 10, keep_u: bool = False):
    """Run one form of the step on workload w: `warmup` untimed steps,
    `steps` timed steps (no instrumentation), then `isteps` steps with HIP
    events around every launch for the kernel table.  Returns steps/s,
    seconds, per-family (total ms, launches), the last CFL*dt and, with
    keep_u, the final velocities.
    model: "" (DNS), "adm-rt" (the relaxation model of sgs.i iles = 1 with
    the default sgs.i), "smag" (the dynamic Smagorinsky model of sgs.i
    iles = 2, cs = 0, ineg = 1) or "learned" (the closure network evaluated at every
    point of the dealiased grid every substep; its cost does not depend on
    the weights, so the network keeps its seeded initial weights)."""
    import numpy as np

    from simson_amd.gpu import BlaGPU

    f = w.field()
    g = BlaGPU(w.nx, w.ny, w.nz, w.re, w.xl, w.zl, cflux=True, precision=precision)
    try:
        if model == "adm-rt":
            g.les(iles=1, cutoff=2.0 / 3.0, iord=5, chi=0.2)
        elif model == "smag":
            g.les(iles=2, cs=0.0, ineg=1)
        elif model == "learned":
            from simson_amd.sgsnet import SgsNet
            with SgsNet(1) as net:
                params = net.params
            utau2 = (180.0 / w.re) ** 2
            y = np.cos(np.pi * np.arange(w.ny) / (w.ny - 1))
            g.closure(params, np.full(9, 1.0 / (utau2 * w.re)), np.zeros(9),
                      (1.0 - np.abs(y)) * 180.0 / 100.0, np.full(6, 0.1 * utau2), np.zeros(6))
        g.upload_velocity(f.u, -2.0 / w.re)
        g.run(warmup, w.dt, icfl=3, first=True)
        g.sync()
        t0 = time.perf_counter()
        cfl = g.run(steps, w.dt, icfl=3, first=False)
        g.sync()
        t1 = time.perf_counter()
        g.timing(True)
        g.timing_read()                     # drop anything recorded so far
        g.run(isteps, w.dt, icfl=3, first=False)
        g.sync()
        kern = g.timing_read()
        g.timing(False)
        u = g.velocities() if keep_u else None
    finally:
        g.close()
    return {"sps": steps / (t1 - t0), "sec": t1 - t0, "kern": kern, "cfl": cfl, "u": u}


def time_spatial(precision: int, steps: int, warmup: int):
    """The spatial Blasius boundary layer forced at F = 100 (512 x 97 x 64,
    Re_delta* 400 at the inflow, box 1000 x 30 x 50 delta*, fringe, ibc 101,
    the lower-wall strip, dt 0.02 delta*/U): steps/s of the step with the
    fringe, the strip's wall velocity and the free-stream boundary rows."""
    from simson_amd.baseflow import tollmien_schlichting_case
    from simson_amd.gpu import BlaGPU

    c = tollmien_schlichting_case()
    f = c.field
    g = BlaGPU(f.nx, f.nyp, f.nz, f.re, f.xl, f.zl, cflux=False, retau=0.0,
               precision=precision, ibc=101)
    try:
        g.fringe(*c.fringe, base=c.base)
        g.free_stream(c.bj, c.du0upp)
        g.wall(*c.wall)
        g.upload_velocity(f.u, 0.0)
        g.run(warmup, c.dt, icfl=3, first=True)
        g.sync()
        t0 = time.perf_counter()
        cfl = g.run(steps, c.dt, icfl=3, first=False)
        g.sync()
        t1 = time.perf_counter()
    finally:
        g.close()
    sps = steps / (t1 - t0)
    return {"steps_per_s": round(sps, 2), "ms_per_step": round(1e3 / sps, 4),
            "grid_points_steps_per_s": round(sps * f.nx * f.nyp * f.nz, 1), "cfl_dt": cfl}


def time_entry_points(reps: int = 5):
    """Milliseconds per call of the FP64 statistics and pressure entry points
    on the headline workload (boxamp, boxxys 1-15, nonlinp + linearp), each
    call synchronising; not part of the step."""
    import numpy as np

    from simson_amd.gpu import SIMSON_FP64, BlaGPU

    w = CHANNEL
    f = w.field()
    g = BlaGPU(w.nx, w.ny, w.nz, w.re, w.xl, w.zl, cflux=True, precision=SIMSON_FP64)
    out = {}
    try:
        g.upload_velocity(f.u, -2.0 / w.re)
        g.run(1, w.dt, icfl=0, first=True)
        acc = np.zeros((15, w.ny, w.nx))
        for name, call in (("amp", g.amp), ("xys", lambda: g.xys(w.dt, acc)),
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


def kernel_table(kern: dict, mixed: bool, w: Workload = CHANNEL):
    alg = alg_per_launch(mixed, w)
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
    # step 2 as a whole: the three kernels against the 5 + 3 fields
    if all(k in kern and kern[k][1] for k in ("zsyn", "xrow", "zana")):
        us = sum(1e3 * kern[k][0] / kern[k][1] for k in ("zsyn", "xrow", "zana"))
        gbs = step2_bytes(mixed, w) / (us * 1e-6) / 1e9
        kernels["step2"] = {"avg_us": round(us, 3), "bound": "hbm",
                            "alg_bytes": step2_bytes(mixed, w), "achieved_GBs": round(gbs, 1),
                            "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return kernels


def roofline_of(kernels: dict, kern: dict, mixed: bool, w: Workload = CHANNEL):
    alg = alg_per_launch(mixed, w)
    dom = max((k for k in kern if k in alg and kern[k][1]), key=lambda k: kern[k][0])
    d = kernels[dom]
    traffic = profiled_traffic(dom, mixed) if w == CHANNEL else None
    if d["bound"] == "hbm":
        out = {"kernel": dom, "bound": "hbm", "achieved": d["achieved_GBs"],
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"], "traffic": traffic}
        if traffic:
            out["traffic_vs_alg"] = round(traffic / d["alg_bytes"], 3)
        return out
    return {"kernel": dom, "bound": "mfma", "achieved": d["achieved_TFLOPs"],
            "peak": FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": d["frac"], "traffic": traffic}


def form_entry(r: dict, mixed: bool, w: Workload):
    sps = r["sps"]
    sb = step_bytes(mixed, w)
    return {"steps_per_s": round(sps, 3), "ms_per_step": round(1e3 / sps, 4),
            "grid_points_steps_per_s": round(sps * w.nx * w.ny * w.nz, 1),
            "step_alg_bytes": sb,
            "step_hbm_frac": round(sb * sps / 1e9 / HBM_PEAK_GBS, 4),
            "cfl_dt": r["cfl"], "kernels": kernel_table(r["kern"], mixed, w)}


def profile_error(u64, u32, re):
    """Deviation of the mixed run's mean and rms profiles and wall shear
    from FP64's after the same steps (largest over y, relative to FP64's
    largest value)."""
    from simson_amd.profiles import deviation, profiles, wall_shear
    m64, r64, uv64 = profiles(u64)
    m32, r32, uv32 = profiles(u32)
    t64, t32 = wall_shear(m64[0], re), wall_shear(m32[0], re)
    return {"mean_u": deviation(m32[0], m64[0]), "rms_u": deviation(r32[0], r64[0]),
            "rms_v": deviation(r32[1], r64[1]), "rms_w": deviation(r32[2], r64[2]),
            "uv": deviation(uv32, uv64), "wall_shear": abs(t32 - t64) / abs(t64)}


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
    ap.add_argument("--large-steps", type=int, default=20,
                    help="timed steps of the 384x257x384 channel per form (0: skip)")
    ap.add_argument("--spatial-steps", type=int, default=200,
                    help="timed steps of the 512x97x64 spatial boundary layer per form "
                         "(0: skip)")
    ap.add_argument("--no-long-check", dest="long_check", action="store_false",
                    help="skip the 10^4-step mixed-vs-FP64 statistics check")
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

    def say(msg):
        print(msg, file=sys.stderr, flush=True)

    # the CPU leg runs first, before this process touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_steps)
    say(f"bench: cpu baseline {cpu['value'] if cpu else None}")

    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED

    w = CHANNEL
    res = {}
    if args.form in ("both", "fp64"):
        res["fp64"] = time_form(w, SIMSON_FP64, args.steps, args.warmup, keep_u=True)
        say(f"bench: fp64 {res['fp64']['sps']:.1f} steps/s")
    if args.form in ("both", "mixed"):
        res["mixed"] = time_form(w, SIMSON_MIXED, args.steps, args.warmup, keep_u=True)
        say(f"bench: mixed {res['mixed']['sps']:.1f} steps/s")
        for m in [m for m in args.models.split(",") if m]:
            res["mixed+" + m] = time_form(w, SIMSON_MIXED, max(1, args.steps // 2), args.warmup,
                                          m)
            say(f"bench: mixed+{m} {res['mixed+' + m]['sps']:.1f} steps/s")
    large = {}
    if args.large_steps > 0:
        for name, prec in (("fp64", SIMSON_FP64), ("mixed", SIMSON_MIXED)):
            r = time_form(LARGE, prec, args.large_steps, 2, isteps=3)
            large[name] = form_entry(r, prec == SIMSON_MIXED, LARGE)
            say(f"bench: {LARGE.tag} {name} {r['sps']:.2f} steps/s")
    spatial = {}
    if args.spatial_steps > 0:
        for name, prec in (("fp64", SIMSON_FP64), ("mixed", SIMSON_MIXED)):
            spatial[name] = time_spatial(prec, args.spatial_steps, 5)
            say(f"bench: spatial boundary layer 512x97x64 {name} "
                f"{spatial[name]['steps_per_s']:.1f} steps/s")
    entry = time_entry_points() if args.entry_points else None
    err = None
    if "fp64" in res and "mixed" in res:
        err = profile_error(res["fp64"]["u"], res["mixed"]["u"], w.re)
        err["steps"] = args.warmup + args.steps + 10
    long_check = None
    if args.long_check and args.form == "both":
        from simson_amd.profiles import long_run
        say("bench: 10^4-step mixed vs FP64 statistics")
        long_check = long_run(w.nx, w.ny, w.nz, w.re, w.xl, w.zl, w.dt, spinup=2000,
                              nsteps=10000, every=250, controls=(1, 2), say=say)
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
    r = res[head]
    kernels = kernel_table(r["kern"], mixed)
    forms = {name: form_entry(fr, name.startswith("mixed"), w) for name, fr in res.items()}
    bar = None
    if err is not None:
        worst = max(v for k, v in err.items() if k != "steps")
        bar = {"bar": 1e-5, "instantaneous_after_steps": err["steps"],
               "instantaneous_max": worst, "met_instantaneous": worst <= 1e-5}
        if long_check:
            ok = all(d["mixed"] <= 3.0 * d["control"] for d in long_check["averaged"].values())
            bar["long_run_within_control_spread"] = ok
    out = {
        "metric": METRIC,
        "value": round(r["sps"], 3),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["sec"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.label} {w.tag}, RK3-CN dt={w.dt}, constant flux, "
                               + ("FP32-mixed step" if mixed else "FP64 step"),
                   "nx": w.nx, "ny": w.ny, "nz": w.nz, "re_cl": w.re, "lx": "4pi",
                   "lz": "4pi/3", "dt": w.dt, "icfl": 3, "form": head,
                   "parallelism": "single GPU"},
        "grid_points_steps_per_s": round(r["sps"] * w.nx * w.ny * w.nz, 1),
        "mixed_vs_fp64": (round(res["mixed"]["sps"] / res["fp64"]["sps"], 3)
                          if "mixed" in res and "fp64" in res else None),
        "cfl_dt": r["cfl"],
        "roofline": roofline_of(kernels, r["kern"], mixed),
        "step_hbm_frac": forms[head]["step_hbm_frac"],
        "mixed_error": err,
        "error_bar": bar,
        "long_check": long_check,
        "forms": forms,
        "large_channel": ({"workload": f"{LARGE.label} {LARGE.tag}, Re_cl {LARGE.re:g}, "
                                       f"lx 2pi, lz pi, dt {LARGE.dt}", **large}
                          if large else None),
        "spatial_boundary_layer": ({"workload": "spatial Blasius boundary layer 512x97x64, "
                                                "Re_delta* 400, fringe, ibc 101, wall strip "
                                                "at F = 100, dt 0.02", **spatial}
                                   if spatial else None),
        "entry_points_ms": entry,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
