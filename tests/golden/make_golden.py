"""Generate the golden fixtures from the reference Simson `bla`.

Runs in the build container only (it needs /root/reference and amdflang):

    python tests/golden/make_golden.py            # c1_32x33x32.npz
    python tests/golden/make_golden.py long       # c1_1000_32x33x32.npz
    python tests/golden/make_golden.py pressure   # p_32x33x32.npz
    python tests/golden/make_golden.py les        # les_32x33x32.npz
    python tests/golden/make_golden.py xys        # xys_32x33x32.npz
    python tests/golden/make_golden.py smag       # smag_32x33x32.npz
    python tests/golden/make_golden.py fringe     # fringe_32x33x32.npz
    python tests/golden/make_golden.py wall       # wall_32x33x32.npz
    python tests/golden/make_golden.py vart       # vart_32x33x32.npz
    python tests/golden/make_golden.py blasius    # blasius_64x33x8.npz
    python tests/golden/make_golden.py tblasius   # tblasius_32x33x16.npz
    python tests/golden/make_golden.py locf       # locf_64x33x8.npz

1. builds the reference `bla` for 32x33x32 (serial, pressure off) with
   oracle/build_ref.sh into oracle/_ref/;
2. writes the seeded plane-channel field (Re_cl 3000, lx 4 pi, lz 2 pi,
   |kx|,|kz| <= 8, rms 1e-3, seed 1, Poiseuille mean) with simson_amd.fields;
3. runs 100 RK3 steps at dt = 0.01 under constant flux, saving the velocity
   field after 1, 10 and 100 steps, with amp.stat written every full step
   (its column 12 is the pressure gradient px, wampbl.f:276-278) and the CFL
   evaluated every substep;
4. stores the reference outputs in c1_32x33x32.npz: the three fields
   (complex128 [comp, z, y, kx], exactly as read from bla's .u files), the
   whole amp.stat table (19 columns per full step, wampbl.f:276-278; column
   12 is px), the CFL*dt log lines and the sha256 of the input file.

The fixture holds data produced by running the reference, not its source.
"""
from __future__ import annotations

import hashlib
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from simson_amd import blai, fields, sgsi, stats, ufile  # noqa: E402

NX, NY, NZ = 32, 33, 32
BIN = os.path.join(REPO, "oracle", "_ref", "bla_32x33x32_serial")
PBIN = os.path.join(REPO, "oracle", "_ref", "bla_32x33x32_press")
SBIN = os.path.join(REPO, "oracle", "_ref", "bla_32x33x32_smag")


def c1_field():
    return fields.channel_field(NX, NY, NZ, 4 * math.pi, 2 * math.pi, 3000.0,
                                kmax=8, rms=1e-3, seed=1)


def main():
    if not os.path.exists(BIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "0",
                               os.path.basename(BIN)])
    work = tempfile.mkdtemp(prefix="simson_golden.")
    try:
        ufile.write_u(os.path.join(work, "init.u"), c1_field())
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=300,
                           dt=0.01, nst=3, xl=4 * math.pi, cflux=True, icfl=3,
                           iamp=3, namamp="amp.stat",
                           tsave=[0.01, 0.1], nmsave=["t1.u", "t10.u"])
        with open(os.path.join(work, "bla.i"), "w") as f:
            f.write(blai.format_channel(p))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([BIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla did not terminate normally")
        cfl = [float(m) for m in re.findall(r"CFL\*dt\s+([0-9.Ee+-]+)", log)]
        amp = np.loadtxt(os.path.join(work, "amp.stat"))
        sha = hashlib.sha256(open(os.path.join(work, "init.u"), "rb").read()).hexdigest()
        u1 = ufile.read_u(os.path.join(work, "t1.u")).u
        u10 = ufile.read_u(os.path.join(work, "t10.u")).u
        u100 = ufile.read_u(os.path.join(work, "end.u")).u
        np.savez_compressed(os.path.join(HERE, "c1_32x33x32.npz"),
                            u1=u1, u10=u10, u100=u100,
                            amp_t=amp[:, 0], amp_px=amp[:, 11], amp=amp,
                            cfl=np.array(cfl), init_sha256=np.array(sha))
        print("wrote", os.path.join(HERE, "c1_32x33x32.npz"), "cfl lines", len(cfl),
              "amp rows", amp.shape)
    finally:
        shutil.rmtree(work)


def long():
    """The C1 case over its full 1000 steps (maxit 3000 substeps): the fields
    after 500 and 1000 steps, amp.stat every 10th step and the CFL*dt lines,
    into c1_1000_32x33x32.npz."""
    if not os.path.exists(BIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "0",
                               os.path.basename(BIN)])
    work = tempfile.mkdtemp(prefix="simson_golden_l.")
    try:
        ufile.write_u(os.path.join(work, "init.u"), c1_field())
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=3000,
                           dt=0.01, nst=3, xl=4 * math.pi, cflux=True, icfl=3,
                           iamp=30, namamp="amp.stat", tsave=[5.0], nmsave=["t500.u"])
        with open(os.path.join(work, "bla.i"), "w") as f:
            f.write(blai.format_channel(p))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([BIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla did not terminate normally")
        cfl = [float(m) for m in re.findall(r"CFL\*dt\s+([0-9.Ee+-]+)", log)]
        amp = np.loadtxt(os.path.join(work, "amp.stat"))
        sha = hashlib.sha256(open(os.path.join(work, "init.u"), "rb").read()).hexdigest()
        np.savez_compressed(os.path.join(HERE, "c1_1000_32x33x32.npz"),
                            u500=ufile.read_u(os.path.join(work, "t500.u")).u,
                            u1000=ufile.read_u(os.path.join(work, "end.u")).u,
                            amp=amp, cfl=np.array(cfl), init_sha256=np.array(sha))
        print("wrote", os.path.join(HERE, "c1_1000_32x33x32.npz"), "cfl lines", len(cfl),
              "amp rows", amp.shape)
    finally:
        shutil.rmtree(work)


def pressure():
    """Pressure fixture: the same run with bla built with pressure = 1 (par.f),
    10 steps, saving t1.u/.u.p and t10.u/.u.p; the .u.p files hold the
    pressure slot written by wdiscp.f (linearp.f after nonlinp.f at the last
    substep before each save)."""
    if not os.path.exists(PBIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "1",
                               os.path.basename(PBIN)])
    work = tempfile.mkdtemp(prefix="simson_golden_p.")
    try:
        ufile.write_u(os.path.join(work, "init.u"), c1_field())
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30,
                           dt=0.01, nst=3, xl=4 * math.pi, cflux=True, icfl=3,
                           tsave=[0.01, 0.1], nmsave=["t1.u", "t10.u"])
        with open(os.path.join(work, "bla.i"), "w") as f:
            f.write(blai.format_channel(p))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([PBIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla (pressure) did not terminate normally")
        res = {}
        for n in ("1", "10"):
            res["u" + n] = ufile.read_u(os.path.join(work, f"t{n}.u")).u
            res["p" + n] = ufile.read_u(os.path.join(work, f"t{n}.u.p"), ncomp=1).u[0]
        np.savez_compressed(os.path.join(HERE, "p_32x33x32.npz"), **res)
        print("wrote", os.path.join(HERE, "p_32x33x32.npz"))
    finally:
        shutil.rmtree(work)


def les_field():
    """The LES runs start from a field with every resolved wavenumber
    excited (|kx|, |kz| <= 15, rms 1e-2, seed 5), so that the high-pass
    filtered relaxation term moves v and w by about 1e-2 relative within 10
    steps (with the reference defaults) and parity at 1e-12 pins it."""
    return fields.channel_field(NX, NY, NZ, 4 * math.pi, 2 * math.pi, 3000.0,
                                kmax=15, rms=1e-2, seed=5)


LES_CASES = {
    # sgs.i defaults of rparamles.f (ADM-RT: cutoff_inv 1.5, iord 5, chi 0.2),
    # 100 steps, saves after 1, 10 and 100 steps
    "a": (sgsi.SgsParams(iles=1, cutoff_inv=1.5, iord=5, chi=0.2), 100, [0.01, 0.1]),
    # another filter order and cutoff with a stronger relaxation, 10 steps
    "b": (sgsi.SgsParams(iles=1, cutoff_inv=3.0, iord=2, chi=2.0), 10, []),
}


def les():
    """ADM-RT fixture: bla.i sets sgs, sgs.i selects iles = 1; the reference
    binary is the serial 32x33x32 one (bla.f:27 compiles LES support in).
    Keys: a_u1, a_u10, a_u100 (case a), b_u10 (case b), and the DNS field
    dns_u10 of the same start without the model (it shows the size of the
    model's effect)."""
    if not os.path.exists(BIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "0",
                               os.path.basename(BIN)])
    res = {}
    runs = [(k, sp, n, ts) for k, (sp, n, ts) in LES_CASES.items()] + [("dns", None, 10, [])]
    for key, sp, nsteps, tsave in runs:
        work = tempfile.mkdtemp(prefix="simson_golden_les.")
        try:
            ufile.write_u(os.path.join(work, "init.u"), les_field())
            names = [f"t{round(t / 0.01)}.u" for t in tsave]
            p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=3 * nsteps,
                               dt=0.01, nst=3, xl=4 * math.pi, cflux=True, icfl=3,
                               sgs=sp is not None, tsave=tsave, nmsave=names)
            with open(os.path.join(work, "bla.i"), "w") as f:
                f.write(blai.format_channel(p))
            if sp is not None:
                with open(os.path.join(work, "sgs.i"), "w") as f:
                    f.write(sgsi.format_sgs(sp))
            with open(os.path.join(work, "out.txt"), "w") as out:
                subprocess.check_call([BIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
            log = open(os.path.join(work, "out.txt")).read()
            if "bla terminated normally" not in log:
                raise RuntimeError(f"reference bla (LES {key}) did not terminate normally")
            for name in names:
                res[f"{key}_u{name[1:-2]}"] = ufile.read_u(os.path.join(work, name)).u
            res[f"{key}_u{nsteps}"] = ufile.read_u(os.path.join(work, "end.u")).u
        finally:
            shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, "les_32x33x32.npz"), **res)
    print("wrote", os.path.join(HERE, "les_32x33x32.npz"), sorted(res))


def xys():
    """x-y statistics fixture: the les_field start (every resolved mode
    excited, so that the Reynolds stresses are well above round-off), no
    model, maxit = 30 at dt = 0.01 with ixys = 9 and ixyss = 15.  boxxys runs
    at the start of substeps 4, 13, 22 and 31 (bla.f:1830), i.e. on the
    fields after 1, 4, 7 and 10 full steps; with statistics on, the stop
    criteria are only checked when mod(it-1, ixys) = 0 (bla.f:2789), so the
    run ends after 12 steps, and the file written then (bla.f:2907-2930)
    holds the dt-weighted mean of the four samples (sumw = 0.04).  Keys: xys
    (statistics 1-15, [15][ny][nx] as stored, divided by sumw), sumw, t, uend
    (the end field)."""
    if not os.path.exists(BIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "0",
                               os.path.basename(BIN)])
    work = tempfile.mkdtemp(prefix="simson_golden_xys.")
    try:
        ufile.write_u(os.path.join(work, "init.u"), les_field())
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30,
                           dt=0.01, nst=3, xl=4 * math.pi, cflux=True, icfl=3,
                           ixys=9, namxys="xy.stat", ixyss=15, txys=0.0)
        with open(os.path.join(work, "bla.i"), "w") as f:
            f.write(blai.format_channel(p))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([BIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla (xys) did not terminate normally")
        xf = stats.read_xys(os.path.join(work, "xy.stat"))
        uend = ufile.read_u(os.path.join(work, "end.u")).u
        np.savez_compressed(os.path.join(HERE, "xys_32x33x32.npz"), xys=xf.stats[:15],
                            sumw=xf.sumw, t=xf.t, uend=uend)
        print("wrote", os.path.join(HERE, "xys_32x33x32.npz"), xf.stats.shape, xf.sumw, xf.t)
    finally:
        shutil.rmtree(work)


SMAG_CASES = {
    # dynamic procedure (cs = 0) with each clipping of rparamles.f:36-39, and
    # a fixed coefficient; 10 steps, saves after 1 step for d1
    "d1": (sgsi.SgsParams(iles=2, cs=0.0, ineg=1), [0.01]),
    "d0": (sgsi.SgsParams(iles=2, cs=0.0, ineg=0), []),
    "d2": (sgsi.SgsParams(iles=2, cs=0.0, ineg=2), []),
    "c": (sgsi.SgsParams(iles=2, cs=0.15, ineg=1), []),
}


HPF_CASES = {
    # the Smagorinsky model on the high-pass filtered velocity (iles = 3):
    # a fixed coefficient, and the dynamic procedure with clipping; 10 steps,
    # saves after 1 step for d1
    "c": (sgsi.SgsParams(iles=3, cutoff_inv=1.5, iord=5, cs=0.15, ineg=1), []),
    "d1": (sgsi.SgsParams(iles=3, cutoff_inv=1.5, iord=5, cs=0.0, ineg=1), [0.01]),
    "d0": (sgsi.SgsParams(iles=3, cutoff_inv=3.0, iord=2, cs=0.0, ineg=0), []),
}


def smag(cases=None, outname="smag_32x33x32.npz"):
    """Eddy-viscosity fixture (sgs.i iles = 2, or iles = 3 with
    cases=HPF_CASES): the reference is compiled with iiles_smag = 1
    (oracle/build_ref.sh with SMAG=1), started from the LES field, 10 steps
    at dt = 0.01.  Keys <case>_u10 (and d1_u1)."""
    cases = SMAG_CASES if cases is None else cases
    if not os.path.exists(SBIN):
        env = dict(os.environ, SMAG="1")
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "0",
                               os.path.basename(SBIN)], env=env)
    res = {}
    for key, (sp, tsave) in cases.items():
        work = tempfile.mkdtemp(prefix="simson_golden_smag.")
        try:
            ufile.write_u(os.path.join(work, "init.u"), les_field())
            names = [f"t{round(t / 0.01)}.u" for t in tsave]
            p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30,
                               dt=0.01, nst=3, xl=4 * math.pi, cflux=True, icfl=3,
                               sgs=True, tsave=tsave, nmsave=names)
            with open(os.path.join(work, "bla.i"), "w") as f:
                f.write(blai.format_channel(p))
            with open(os.path.join(work, "sgs.i"), "w") as f:
                f.write(sgsi.format_sgs(sp))
            with open(os.path.join(work, "out.txt"), "w") as out:
                subprocess.check_call([SBIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
            log = open(os.path.join(work, "out.txt")).read()
            if "bla terminated normally" not in log:
                raise RuntimeError(f"reference bla (smag {key}) did not terminate normally")
            for name in names:
                res[f"{key}_u{name[1:-2]}"] = ufile.read_u(os.path.join(work, name)).u
            res[f"{key}_u10"] = ufile.read_u(os.path.join(work, "end.u")).u
        finally:
            shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, outname), **res)
    print("wrote", os.path.join(HERE, outname), sorted(res))


FRINGE_CASES = {
    # (fmax, fstart, fend, frise, ffall) in the channel scaling of bla.i
    "a": (1.0, -4.0, 0.0, 2.0, 1.0),
    "b": (2.5, -6.0, 3.0, 3.0, 2.0),
}


def fringe():
    """Fringe fixture: the spatial Poiseuille channel (fltype 4, spat, ibc 0,
    no pressure gradient: bla.f:1657-1662) from the LES start field, the
    fringe damping toward U = 1 - y^2 (cbflow.f:117-131), 10 steps at dt =
    0.01, saves after 1 and 10 steps with their pressure (the pressure = 1
    binary: fringp.f adds the same forcing to the pressure's nonlinear
    term).  Keys <case>_u1, <case>_u10, <case>_p10."""
    import dataclasses
    if not os.path.exists(PBIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "1",
                               os.path.basename(PBIN)])
    res = {}
    for key, (fmax, fstart, fend, frise, ffall) in FRINGE_CASES.items():
        work = tempfile.mkdtemp(prefix="simson_golden_fringe.")
        try:
            f = dataclasses.replace(les_field(), fltype=4,
                                    extra={"bstart": 0.0, "blength": 0.0})
            ufile.write_u(os.path.join(work, "init.u"), f)
            p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30,
                               dt=0.01, nst=3, xl=4 * math.pi, cflux=False, retau=0.0,
                               icfl=3, spat=True, fmax=fmax, fstart=fstart, fend=fend,
                               frise=frise, ffall=ffall, tsave=[0.01, 0.1],
                               nmsave=["t1.u", "t10.u"])
            with open(os.path.join(work, "bla.i"), "w") as fh:
                fh.write(blai.format_channel(p))
            with open(os.path.join(work, "out.txt"), "w") as out:
                subprocess.check_call([PBIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
            log = open(os.path.join(work, "out.txt")).read()
            if "bla terminated normally" not in log:
                raise RuntimeError(f"reference bla (fringe {key}) did not terminate normally")
            res[f"{key}_u1"] = ufile.read_u(os.path.join(work, "t1.u")).u
            res[f"{key}_u10"] = ufile.read_u(os.path.join(work, "t10.u")).u
            res[f"{key}_p10"] = ufile.read_u(os.path.join(work, "t10.u.p"), ncomp=1).u[0]
        finally:
            shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, "fringe_32x33x32.npz"), **res)
    print("wrote", os.path.join(HERE, "fringe_32x33x32.npz"), sorted(res))


WALL_CASES = {
    # (amp, damp, xstart, xend, xrise, xfall, zbet, tomeg) of bla.i wbci = 1
    "a": (0.05, 0.5, -3.0, 2.0, 1.0, 1.5, 2.0, 20.0),
    "b": (0.02, 1.0, -5.0, -1.0, 2.0, 0.5, 1.0, 0.0),
}


def wall():
    """Blowing and suction fixture: the temporal channel from the LES start
    field with the lower-wall strip of bla.i wbci = 1 (cwallbc.f), 10 steps
    at dt = 0.01 under constant flux; case a oscillates in time with the
    suction part damped, case b is steady.  Keys <case>_u1, <case>_u10 and
    the CFL*dt log lines <case>_cfl (boxcfl.f without the wall cap)."""
    if not os.path.exists(BIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "0",
                               os.path.basename(BIN)])
    res = {}
    for key, (amp, damp, x0, x1, xr, xf, zb, om) in WALL_CASES.items():
        work = tempfile.mkdtemp(prefix="simson_golden_wall.")
        try:
            ufile.write_u(os.path.join(work, "init.u"), les_field())
            p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30,
                               dt=0.01, nst=3, xl=4 * math.pi, cflux=True, icfl=3, wbci=1,
                               wamp=amp, wdamp=damp, wxstart=x0, wxend=x1, wxrise=xr,
                               wxfall=xf, wzbet=zb, wtomeg=om, tsave=[0.01], nmsave=["t1.u"])
            with open(os.path.join(work, "bla.i"), "w") as fh:
                fh.write(blai.format_channel(p))
            with open(os.path.join(work, "out.txt"), "w") as out:
                subprocess.check_call([BIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
            log = open(os.path.join(work, "out.txt")).read()
            if "bla terminated normally" not in log:
                raise RuntimeError(f"reference bla (wall {key}) did not terminate normally")
            res[f"{key}_u1"] = ufile.read_u(os.path.join(work, "t1.u")).u
            res[f"{key}_u10"] = ufile.read_u(os.path.join(work, "end.u")).u
            res[f"{key}_cfl"] = np.array(
                [float(m) for m in re.findall(r"CFL\*dt\s+([0-9.Ee+-]+)", log)])
        finally:
            shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, "wall_32x33x32.npz"), **res)
    print("wrote", os.path.join(HERE, "wall_32x33x32.npz"), sorted(res))


def vart():
    """Variable time step fixture (bla.i dt = 0, rparambl.f:1996-2017): the
    C1 field, cflmaxin 0.8, 20 steps, a save at t = 1 (which makes bla.f
    shorten the steps before it, bla.f:2355-2380), the CFL evaluated every
    full step.  Keys: u_save (the field at t = 1), u_end (after 20 steps), cfl
    and dt (every CFL*dt log line: the CFL number and the step)."""
    if not os.path.exists(BIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(NX), str(NY), str(NZ), "1", "no", "0",
                               os.path.basename(BIN)])
    work = tempfile.mkdtemp(prefix="simson_golden_v.")
    try:
        ufile.write_u(os.path.join(work, "init.u"), c1_field())
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60,
                           dt=0.0, nst=3, cflmaxin=0.8, xl=4 * math.pi, cflux=True, icfl=3,
                           tsave=[1.0], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as f:
            f.write(blai.format_channel(p))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([BIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla (vart) did not terminate normally")
        pairs = re.findall(r"CFL\*dt\s+([0-9.Ee+-]+)\s+dt\s+([0-9.Ee+-]+)", log)
        first = re.search(r"First time step\s*:\s*([0-9.Ee+-]+)", log)
        sha = hashlib.sha256(open(os.path.join(work, "init.u"), "rb").read()).hexdigest()
        np.savez_compressed(os.path.join(HERE, "vart_32x33x32.npz"),
                            u_save=ufile.read_u(os.path.join(work, "t1.u")).u,
                            t_save=np.array(ufile.read_u(os.path.join(work, "t1.u")).t),
                            u_end=ufile.read_u(os.path.join(work, "end.u")).u,
                            t_end=np.array(ufile.read_u(os.path.join(work, "end.u")).t),
                            cfl=np.array([float(a) for a, _ in pairs]),
                            dt=np.array([float(b) for _, b in pairs]),
                            dt_first=np.array(float(first.group(1))),
                            init_sha256=np.array(sha))
        print("wrote", os.path.join(HERE, "vart_32x33x32.npz"), "lines", len(pairs))
    finally:
        shutil.rmtree(work)


# spatial Blasius boundary layer at 64x33x8 (all lengths in displacement
# thicknesses of the box centre, as bla.i gives them; the .u file holds them
# times dstar = 2/yl)
BL = dict(nx=64, ny=33, nz=8, re=400.0, xl=100.0, yl=20.0, zl=20.0, bstart=-35.0,
          blength=25.0, dt=0.1, kmax=4, rms=1e-2, seed=5,
          fringe=(1.0, -45.0, 0.0, 15.0, 10.0))
BL_WALL = (0.01, 1.0, -20.0, -10.0, 3.0, 3.0, 0.5, 0.3)
BLBIN = os.path.join(REPO, "oracle", "_ref", "bla_64x33x8_serial")


def fsc_dat(work):
    """Run the reference similarity solver (oracle/_ref/fsc, built by
    oracle/build_fsc.sh) with bls/fsc.i's defaults for Blasius: m 0,
    n 20001, eps 1e-15, ymax 30, no scalar, no text output."""
    exe = os.path.join(REPO, "oracle", "_ref", "fsc")
    if not os.path.exists(exe):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_fsc.sh")])
    with open(os.path.join(work, "fsc.i"), "w") as f:
        f.write("0\n20001\n1e-15\n30\n0\n0\n")
    with open(os.path.join(work, "fsc.out"), "w") as out:
        subprocess.check_call([exe], cwd=work, stdout=out, stderr=subprocess.STDOUT)
    return os.path.join(work, "fsc.dat")


def blasius_field(table):
    from simson_amd import baseflow
    c = BL
    ds = 2.0 / c["yl"]
    base, _, _ = baseflow.spatial_base_flow(c["nx"], c["ny"], c["xl"] * ds, c["re"] / ds, ds,
                                            table, c["bstart"] * ds, c["blength"] * ds)
    return fields.boundary_layer_field(c["nx"], c["ny"], c["nz"], c["xl"] * ds, c["zl"] * ds,
                                       c["re"] / ds, ds, base, c["bstart"] * ds,
                                       c["blength"] * ds, kmax=c["kmax"], rms=c["rms"],
                                       seed=c["seed"])


def blasius():
    """Spatial Blasius boundary-layer fixture (fltype 6, spat, ibc 101,
    cflux off, no pressure gradient) at 64x33x8, Re_delta* 400, box
    100 x 20 x 20 delta*: the fsc.dat of the reference similarity solver,
    the base flow blended over [-35, -10] delta*, the fringe (fmax 1 on
    [-45, 0], rise 15, fall 10) and a seeded one-wall perturbation (rms
    1e-2, |kx|, |kz| <= 4) that is O(1e-2) up to the free stream, so that
    the free-stream rows matter.  20 steps at dt = 0.1 delta*/U, saves after
    1 and 20 steps; case b adds the lower-wall strip (wbci = 1, amp 0.01 on
    [-20, -10], rise/fall 3, zbet 0.5, tomeg 0.3).  Keys: table (the
    similarity table every 20th row), dstar2, init, <case>_u1, <case>_u20,
    and the log's du0upp (times dstar) and x0 (over dstar)."""
    from simson_amd import baseflow
    c = BL
    if not os.path.exists(BLBIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(c["nx"]), str(c["ny"]), str(c["nz"]), "1", "no", "0",
                               os.path.basename(BLBIN)])
    res = {}
    for key in ("a", "b"):
        work = tempfile.mkdtemp(prefix="simson_golden_bl.")
        try:
            table = baseflow.read_fsc(fsc_dat(work))
            if key == "a":
                res["table"] = table.tab[::20]
                res["dstar2"] = np.array(table.dstar2)
            f = blasius_field(table)
            ufile.write_u(os.path.join(work, "init.u"), f)
            if key == "a":
                res["init"] = ufile.read_u(os.path.join(work, "init.u")).u
            fmax, fstart, fend, frise, ffall = c["fringe"]
            kw = dict(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60, dt=c["dt"],
                      nst=3, xl=c["xl"], cflux=False, retau=0.0, ibc=101, icfl=3, spat=True,
                      fmax=fmax, fstart=fstart, fend=fend, frise=frise, ffall=ffall,
                      tsave=[c["dt"]], nmsave=["t1.u"])
            if key == "b":
                amp, damp, x0, x1, xr, xf, zb, om = BL_WALL
                kw.update(wbci=1, wamp=amp, wdamp=damp, wxstart=x0, wxend=x1, wxrise=xr,
                          wxfall=xf, wzbet=zb, wtomeg=om)
            with open(os.path.join(work, "bla.i"), "w") as fh:
                fh.write(blai.format_channel(blai.BlaParams(**kw)))
            with open(os.path.join(work, "out.txt"), "w") as out:
                subprocess.check_call([BLBIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
            log = open(os.path.join(work, "out.txt")).read()
            if "bla terminated normally" not in log:
                raise RuntimeError(f"reference bla (blasius {key}) did not terminate normally")
            res[f"{key}_u1"] = ufile.read_u(os.path.join(work, "t1.u")).u
            res[f"{key}_u20"] = ufile.read_u(os.path.join(work, "end.u")).u
            if key == "a":
                res["du0upp_log"] = np.array(float(re.search(
                    r"du0upp\s*:\s*([0-9.Ee+-]+)", log).group(1)))
                res["x0_log"] = np.array(float(re.search(
                    r"x0\s*:\s*([0-9.Ee+-]+)", log).group(1)))
        finally:
            shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, "blasius_64x33x8.npz"), **res)
    print("wrote", os.path.join(HERE, "blasius_64x33x8.npz"), sorted(res))


# loctyp 3 volume force (locf.f) on the spatial Blasius case of BL, in
# displacement thicknesses: amplitudes (x, y, z), x scale, x position,
# y scale, x type, frequency
LOCF = ((0.0, 0.01, 0.005), 8.0, -15.0, 2.0, 0.0, 0.3)


def locf():
    """Localized volume force fixture: the spatial Blasius case of blasius()
    (case a: fringe, ibc 101, no wall strip) with bla.i loctyp = 3
    (locf.f:358-410): amplitudes (0, 0.01, 0.005), x scale 8, at x = -15,
    y scale 2, x type 0, frequency 0.3, all in delta* units; 20 steps,
    saves after 1 and 20.  Keys u1, u20 (the start field is blasius_64x33x8's
    init)."""
    from simson_amd import baseflow
    c = BL
    if not os.path.exists(BLBIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(c["nx"]), str(c["ny"]), str(c["nz"]), "1", "no", "0",
                               os.path.basename(BLBIN)])
    res = {}
    work = tempfile.mkdtemp(prefix="simson_golden_locf.")
    try:
        table = baseflow.read_fsc(fsc_dat(work))
        f = blasius_field(table)
        f.u = np.load(os.path.join(HERE, "blasius_64x33x8.npz"))["init"]
        ufile.write_u(os.path.join(work, "init.u"), f)
        fmax, fstart, fend, frise, ffall = c["fringe"]
        amp, xs, x0, ys, xt, om = LOCF
        kw = dict(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60, dt=c["dt"], nst=3,
                  xl=c["xl"], cflux=False, retau=0.0, ibc=101, icfl=3, spat=True, fmax=fmax,
                  fstart=fstart, fend=fend, frise=frise, ffall=ffall, loctyp=3, lamp=amp,
                  lxscale=xs, lxloc0=x0, lyscale=ys, lxtype=xt, ltomeg=om,
                  tsave=[c["dt"]], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(blai.BlaParams(**kw)))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([BLBIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla (locf) did not terminate normally")
        res["u1"] = ufile.read_u(os.path.join(work, "t1.u")).u
        res["u20"] = ufile.read_u(os.path.join(work, "end.u")).u
    finally:
        shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, "locf_64x33x8.npz"), **res)
    print("wrote", os.path.join(HERE, "locf_64x33x8.npz"), sorted(res))


def blasius100():
    """The spatial Blasius case a of blasius() with ibc = 100 (the base
    flow's values, not its derivatives, imposed at y = +1: evalbc.f:142-159
    with bu1jr(:, i, 1)); 20 steps, saves after 1 and 20.  Keys u1, u20 (the
    start field is blasius_64x33x8's init)."""
    from simson_amd import baseflow
    c = BL
    if not os.path.exists(BLBIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(c["nx"]), str(c["ny"]), str(c["nz"]), "1", "no", "0",
                               os.path.basename(BLBIN)])
    res = {}
    work = tempfile.mkdtemp(prefix="simson_golden_bl100.")
    try:
        table = baseflow.read_fsc(fsc_dat(work))
        f = blasius_field(table)
        f.u = np.load(os.path.join(HERE, "blasius_64x33x8.npz"))["init"]
        ufile.write_u(os.path.join(work, "init.u"), f)
        fmax, fstart, fend, frise, ffall = c["fringe"]
        kw = dict(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60, dt=c["dt"], nst=3,
                  xl=c["xl"], cflux=False, retau=0.0, ibc=100, icfl=3, spat=True, fmax=fmax,
                  fstart=fstart, fend=fend, frise=frise, ffall=ffall,
                  tsave=[c["dt"]], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(blai.BlaParams(**kw)))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([BLBIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla (ibc 100) did not terminate normally")
        res["u1"] = ufile.read_u(os.path.join(work, "t1.u")).u
        res["u20"] = ufile.read_u(os.path.join(work, "end.u")).u
    finally:
        shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, "blasius100_64x33x8.npz"), **res)
    print("wrote", os.path.join(HERE, "blasius100_64x33x8.npz"), sorted(res))


TBL = dict(nx=32, ny=33, nz=16, re=400.0, xl=60.0, yl=20.0, zl=30.0, dt=0.1, kmax=4,
           rms=1e-2, seed=8)
TBLBIN = os.path.join(REPO, "oracle", "_ref", "bla_32x33x16_serial")


def tblasius_field(table):
    from simson_amd import baseflow
    c = TBL
    ds = 2.0 / c["yl"]
    base, _, _, _ = baseflow.temporal_base_flow(c["nx"], c["ny"], c["xl"] * ds, c["re"] / ds,
                                                ds, table)
    f = fields.boundary_layer_field(c["nx"], c["ny"], c["nz"], c["xl"] * ds, c["zl"] * ds,
                                    c["re"] / ds, ds, base, 0.0, 0.0, kmax=c["kmax"],
                                    rms=c["rms"], seed=c["seed"], fltype=3)
    f.extra = {}
    return f


def tblasius():
    """Temporal Blasius boundary-layer fixture (fltype 3, not spat, cdev 0,
    ibc 101, no pressure gradient) at 32x33x16, Re_delta* 400, box 60 x 20
    x 30 delta*: fsc.dat of the reference similarity solver, the parallel
    Blasius profile plus a seeded one-wall perturbation (rms 1e-2, |kx|,
    |kz| <= 4); 20 steps at dt = 0.1 delta*/U with the parallel-flow forcing
    of nonlinbl.f:771-795, saves after 1 and 20 steps.  Keys: init, u1,
    u20, and the log's du0upp (times dstar) and x0 (over dstar)."""
    from simson_amd import baseflow
    c = TBL
    if not os.path.exists(TBLBIN):
        subprocess.check_call([os.path.join(REPO, "oracle", "build_ref.sh"),
                               str(c["nx"]), str(c["ny"]), str(c["nz"]), "1", "no", "0",
                               os.path.basename(TBLBIN)])
    res = {}
    work = tempfile.mkdtemp(prefix="simson_golden_tbl.")
    try:
        table = baseflow.read_fsc(fsc_dat(work))
        f = tblasius_field(table)
        ufile.write_u(os.path.join(work, "init.u"), f)
        res["init"] = ufile.read_u(os.path.join(work, "init.u")).u
        kw = dict(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60, dt=c["dt"], nst=3,
                  xl=c["xl"], cflux=False, retau=0.0, ibc=101, icfl=3, spat=False, cdev=0.0,
                  tsave=[c["dt"]], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(blai.BlaParams(**kw)))
        with open(os.path.join(work, "out.txt"), "w") as out:
            subprocess.check_call([TBLBIN], cwd=work, stdout=out, stderr=subprocess.STDOUT)
        log = open(os.path.join(work, "out.txt")).read()
        if "bla terminated normally" not in log:
            raise RuntimeError("reference bla (temporal Blasius) did not terminate normally")
        res["u1"] = ufile.read_u(os.path.join(work, "t1.u")).u
        res["u20"] = ufile.read_u(os.path.join(work, "end.u")).u
        res["du0upp_log"] = np.array(float(re.search(r"du0upp\s*:\s*([0-9.Ee+-]+)",
                                                     log).group(1)))
        res["x0_log"] = np.array(float(re.search(r"x0\s*:\s*([0-9.Ee+-]+)", log).group(1)))
    finally:
        shutil.rmtree(work)
    np.savez_compressed(os.path.join(HERE, "tblasius_32x33x16.npz"), **res)
    print("wrote", os.path.join(HERE, "tblasius_32x33x16.npz"), sorted(res))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "pressure":
        pressure()
    elif len(sys.argv) > 1 and sys.argv[1] == "les":
        les()
    elif len(sys.argv) > 1 and sys.argv[1] == "xys":
        xys()
    elif len(sys.argv) > 1 and sys.argv[1] == "smag":
        smag()
    elif len(sys.argv) > 1 and sys.argv[1] == "hpf":
        smag(HPF_CASES, "hpf_32x33x32.npz")
    elif len(sys.argv) > 1 and sys.argv[1] == "fringe":
        fringe()
    elif len(sys.argv) > 1 and sys.argv[1] == "wall":
        wall()
    elif len(sys.argv) > 1 and sys.argv[1] == "long":
        long()
    elif len(sys.argv) > 1 and sys.argv[1] == "blasius":
        blasius()
    elif len(sys.argv) > 1 and sys.argv[1] == "tblasius":
        tblasius()
    elif len(sys.argv) > 1 and sys.argv[1] == "locf":
        locf()
    elif len(sys.argv) > 1 and sys.argv[1] == "blasius100":
        blasius100()
    elif len(sys.argv) > 1 and sys.argv[1] == "vart":
        vart()
    else:
        main()
