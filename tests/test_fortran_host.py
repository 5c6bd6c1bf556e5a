"""The Fortran host (simson_amd/_lib/blagpu) run like the reference bla.

blagpu reads bla.i and the .u start file, drives the RK3 substeps through
the ISO_C_BINDING interface and writes .u files.  On the GPU it is run on
the golden 32x33x32 case and its t1.u / t10.u / end.u and CFL log lines are
compared with what the reference bla wrote for the same inputs, to the
bounds of the library's own golden tests, and so is its amp.stat file; on
the CPU only its input checks run (they stop before any device call).
"""
import math
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT
from simson_amd import blai, fields, sgsi, ufile

BLAGPU = os.path.join(ROOT, "simson_amd", "_lib", "blagpu")


def _need_binary():
    if not os.path.exists(BLAGPU):
        pytest.fail("blagpu is not built (python -m simson_amd.build)")


def _c1_dir(**over):
    work = tempfile.mkdtemp(prefix="blagpu.")
    f = fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0,
                             kmax=8, rms=1e-3, seed=1)
    ufile.write_u(os.path.join(work, "init.u"), f)
    kw = dict(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=300, dt=0.01, nst=3,
              xl=4 * math.pi, cflux=True, icfl=3, iamp=3, namamp="amp.stat",
              tsave=[0.01, 0.1], nmsave=["t1.u", "t10.u"])
    kw.update(over)
    with open(os.path.join(work, "bla.i"), "w") as fh:
        fh.write(blai.format_channel(blai.BlaParams(**kw)))
    return work


def _run(work, timeout=120):
    return subprocess.run([BLAGPU], cwd=work, capture_output=True, text=True,
                          timeout=timeout)


def test_rejects_wrong_version():
    _need_binary()
    work = _c1_dir()
    try:
        path = os.path.join(work, "bla.i")
        text = open(path).read().replace("20070716", "20070715", 1)
        open(path, "w").write(text)
        r = _run(work)
        assert r.returncode != 0
        assert "345431" in r.stdout
    finally:
        shutil.rmtree(work)


def test_rejects_features_outside_the_library():
    _need_binary()
    work = _c1_dir()
    try:
        path = os.path.join(work, "bla.i")
        # switch on constant wall suction (not part of the step library)
        lines = open(path).read().splitlines()
        k = next(i for i, l in enumerate(lines) if "(suction)" in l)
        lines[k] = ".true.  Constant wall-suction (suction)"
        open(path, "w").write("\n".join(lines) + "\n")
        r = _run(work)
        assert r.returncode != 0
        assert "outside the channel step library" in r.stdout
    finally:
        shutil.rmtree(work)


def _les_dir(sp, **over):
    """LES start field of tests/golden/les_32x33x32.npz, bla.i with sgs set
    and the given sgs.i."""
    work = tempfile.mkdtemp(prefix="blagpu_les.")
    f = fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0,
                             kmax=15, rms=1e-2, seed=5)
    ufile.write_u(os.path.join(work, "init.u"), f)
    kw = dict(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30, dt=0.01, nst=3,
              xl=4 * math.pi, cflux=True, icfl=3, sgs=True, tsave=[0.01], nmsave=["t1.u"])
    kw.update(over)
    with open(os.path.join(work, "bla.i"), "w") as fh:
        fh.write(blai.format_channel(blai.BlaParams(**kw)))
    with open(os.path.join(work, "sgs.i"), "w") as fh:
        fh.write(sp if isinstance(sp, str) else sgsi.format_sgs(sp))
    return work


def test_rejects_unsupported_sgs_models():
    """sgs.i is read before any device call: an unknown model (iles 4) and
    out-of-range parameters (ineg 3, iord 7) stop with rparamles.f's code
    45543, a wrong version with 345432."""
    _need_binary()
    for text, code in (("20060909\n4\n", "45543"),
                       ("20060909\n3\n1.5\n7\n0.1\n", "45543"),
                       ("20060909\n2\n0.\n3\n", "45543"),
                       ("20060909\n1\n1.5\n7\n0.2\n", "45543"),
                       ("20060101\n1\n", "345432")):
        work = _les_dir(text)
        try:
            r = _run(work)
            assert r.returncode != 0
            assert code in r.stdout, r.stdout
        finally:
            shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_les_matches_reference_bla():
    """The host with sgs.i (ADM-RT, the defaults of rparamles.f) writes the
    fields the reference bla wrote (tests/golden/les_32x33x32.npz)."""
    _need_binary()
    print("[gpu] blagpu with sgs.i (ADM-RT) 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "les_32x33x32.npz"))
    work = _les_dir(sgsi.SgsParams(iles=1, cutoff_inv=1.5, iord=5, chi=0.2))
    try:
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u", "a_u1"), ("end.u", "a_u10")):
            u = np.delete(ufile.read_u(os.path.join(work, name)).u, 16, axis=1)
            rr = np.delete(golden[key], 16, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_fortran_host_matches_reference_bla(golden_c1):
    _need_binary()
    work = _c1_dir()
    try:
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "blagpu terminated normally" in r.stdout
        for name, key in (("t1.u", "u1"), ("t10.u", "u10"), ("end.u", "u100")):
            uf = ufile.read_u(os.path.join(work, name))
            assert (uf.nx, uf.nyp, uf.nz) == (32, 33, 32)
            ref = golden_c1[key]
            # the z-oddball plane kz = -nz/2 is left holding the nonlinear
            # term by both codes (linearbl skips it); it is compared apart
            u = np.delete(uf.u, 16, axis=1)
            rr = np.delete(ref, 16, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                assert err <= 1e-12, (name, c, err)
            odd = np.abs(uf.u[:, 16] - ref[:, 16]).max() / np.abs(ref[:, 16]).max()
            assert odd < 1e-8, (name, odd)
        end = ufile.read_u(os.path.join(work, "end.u"))
        assert abs(end.t - 1.0) < 1e-12
        cfl = [float(m) for m in re.findall(r"CFL\*dt\s+([0-9.Ee+-]+)", r.stdout)]
        assert len(cfl) == len(golden_c1["cfl"])
        np.testing.assert_allclose(cfl, golden_c1["cfl"], rtol=1e-12)
        # amp.stat: every line of the reference's, same columns
        amp = np.loadtxt(os.path.join(work, "amp.stat"))
        ref = golden_c1["amp"]
        assert amp.shape == ref.shape
        np.testing.assert_allclose(amp[:, 0], ref[:, 0], rtol=1e-14)
        np.testing.assert_allclose(amp[:, 1:18], ref[:, 1:18], rtol=1e-9)
        np.testing.assert_allclose(amp[:, 18], ref[:, 18], atol=1e-14)
    finally:
        shutil.rmtree(work)


def _xys_dir():
    """The x-y statistics run of tests/golden/xys_32x33x32.npz: the LES
    start field, no model, maxit 30, ixys 9, ixyss 15."""
    work = tempfile.mkdtemp(prefix="blagpu_xys.")
    f = fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0,
                             kmax=15, rms=1e-2, seed=5)
    ufile.write_u(os.path.join(work, "init.u"), f)
    p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30, dt=0.01, nst=3,
                       xl=4 * math.pi, cflux=True, icfl=3, ixys=9, namxys="xy.stat",
                       ixyss=15, txys=0.0)
    with open(os.path.join(work, "bla.i"), "w") as fh:
        fh.write(blai.format_channel(p))
    return work


def test_rejects_correlations():
    """corrf (two-point correlations) is read and refused before any device
    call."""
    _need_binary()
    work = _xys_dir()
    try:
        path = os.path.join(work, "bla.i")
        text = open(path).read().replace(".false.         Two-point correlations in z",
                                         ".true.          Two-point correlations in z")
        open(path, "w").write(text)
        r = _run(work)
        assert r.returncode != 0
        assert "correlations" in r.stdout
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_xys_matches_reference_bla():
    """The host's statistics schedule (samples after 1, 4, 7, 10 steps, the
    stop deferred to a statistics step, i.e. after 12 steps) and its wxys.f
    file match the reference's: statistics 1-15 to the bounds of
    tests/test_xys.py, the end field to 1e-12."""
    from simson_amd import stats
    _need_binary()
    print("[gpu] blagpu with ixys 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "xys_32x33x32.npz"))
    work = _xys_dir()
    try:
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        xf = stats.read_xys(os.path.join(work, "xy.stat"))
        assert (xf.nx, xf.ny, xf.nz) == (32, 33, 32)
        assert abs(xf.sumw - float(golden["sumw"])) < 1e-15
        assert abs(xf.t - float(golden["t"])) < 1e-12
        ref = golden["xys"]
        for i in range(15):
            tol = 1e-10 if i in (6, 8, 9, 11) else 1e-12
            err = np.abs(xf.stats[i] - ref[i]).max() / np.abs(ref[i]).max()
            assert err < tol, (stats.XYS_NAMES[i], err)
        u = np.delete(ufile.read_u(os.path.join(work, "end.u")).u, 16, axis=1)
        rr = np.delete(golden["uend"], 16, axis=1)
        for c in range(3):
            assert _rel(u[c], rr[c]) <= 1e-12, c
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_smag_matches_reference_bla():
    """The host with sgs.i iles = 2 (dynamic Smagorinsky, ineg = 1) writes
    the fields the reference bla compiled with the eddy-viscosity models
    wrote (tests/golden/smag_32x33x32.npz)."""
    _need_binary()
    print("[gpu] blagpu with sgs.i (dynamic Smagorinsky) 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "smag_32x33x32.npz"))
    work = _les_dir(sgsi.SgsParams(iles=2, cs=0.0, ineg=1))
    try:
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u", "d1_u1"), ("end.u", "d1_u10")):
            u = np.delete(ufile.read_u(os.path.join(work, name)).u, 16, axis=1)
            rr = np.delete(golden[key], 16, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_hpf_matches_reference_bla():
    """The host with sgs.i iles = 3 (the dynamic HPF model, ineg = 1) writes
    the fields the reference bla wrote (tests/golden/hpf_32x33x32.npz)."""
    _need_binary()
    print("[gpu] blagpu with sgs.i (dynamic HPF Smagorinsky) 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "hpf_32x33x32.npz"))
    work = _les_dir(sgsi.SgsParams(iles=3, cutoff_inv=1.5, iord=5, cs=0.0, ineg=1))
    try:
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u", "d1_u1"), ("end.u", "d1_u10")):
            u = np.delete(ufile.read_u(os.path.join(work, name)).u, 16, axis=1)
            rr = np.delete(golden[key], 16, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_fringe_matches_reference_bla():
    """The host on a spatial Poiseuille channel (fltype 4, bla.i spat with
    the fringe of case a of tests/golden/fringe_32x33x32.npz) writes the
    fields the reference bla wrote, fltype 4 header included."""
    import dataclasses
    _need_binary()
    print("[gpu] blagpu spatial channel with fringe 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "fringe_32x33x32.npz"))
    work = tempfile.mkdtemp(prefix="blagpu_fringe.")
    try:
        f = fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0,
                                 kmax=15, rms=1e-2, seed=5)
        f = dataclasses.replace(f, fltype=4, extra={"bstart": 0.0, "blength": 0.0})
        ufile.write_u(os.path.join(work, "init.u"), f)
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30, dt=0.01,
                           nst=3, xl=4 * math.pi, cflux=False, retau=0.0, icfl=3, spat=True,
                           fmax=1.0, fstart=-4.0, fend=0.0, frise=2.0, ffall=1.0,
                           tsave=[0.01], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(p))
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u", "a_u1"), ("end.u", "a_u10")):
            uf = ufile.read_u(os.path.join(work, name))
            assert uf.fltype == 4
            u = np.delete(uf.u, 16, axis=1)
            rr = np.delete(golden[key], 16, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_wall_matches_reference_bla():
    """The host with the lower-wall strip (bla.i wbci = 1, case a of
    tests/golden/wall_32x33x32.npz: oscillating, suction damped) writes the
    fields and CFL*dt lines the reference bla wrote."""
    _need_binary()
    print("[gpu] blagpu with wbci = 1 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "wall_32x33x32.npz"))
    work = tempfile.mkdtemp(prefix="blagpu_wall.")
    try:
        f = fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0,
                                 kmax=15, rms=1e-2, seed=5)
        ufile.write_u(os.path.join(work, "init.u"), f)
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=30, dt=0.01,
                           nst=3, xl=4 * math.pi, cflux=True, icfl=3, wbci=1, wamp=0.05,
                           wdamp=0.5, wxstart=-3.0, wxend=2.0, wxrise=1.0, wxfall=1.5,
                           wzbet=2.0, wtomeg=20.0, tsave=[0.01], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(p))
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u", "a_u1"), ("end.u", "a_u10")):
            u = np.delete(ufile.read_u(os.path.join(work, name)).u, 16, axis=1)
            rr = np.delete(golden[key], 16, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
        cfl = [float(m) for m in re.findall(r"CFL\*dt\s+([0-9.Ee+-]+)", r.stdout)]
        np.testing.assert_allclose(cfl, golden["a_cfl"], rtol=1e-10)
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_variable_dt_matches_reference_bla():
    """bla.i dt = 0 (variable time step, rparambl.f:1996-2017): the first
    step from the CFL of the initial field (getdt.f), then each step from
    the latest CFL number, shortened before the save at t = 1
    (bla.f:2355-2380).  The host's dt sequence, CFL lines and fields equal
    the reference run's (tests/golden/vart_32x33x32.npz): dt to 1e-12
    relative (it is a rational function of the CFL maximum), fields 1e-12."""
    _need_binary()
    print("[gpu] blagpu with a variable time step 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "vart_32x33x32.npz"))
    work = _c1_dir(maxit=60, dt=0.0, cflmaxin=0.8, iamp=0, tsave=[1.0], nmsave=["t1.u"])
    try:
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        pairs = re.findall(r"CFL\*dt\s+([0-9.Ee+-]+)\s+dt\s+([0-9.Ee+-]+)", r.stdout)
        dt = np.array([float(b) for _, b in pairs])
        cfl = np.array([float(a) for a, _ in pairs])
        first = float(re.search(r"First time step\s*:\s*([0-9.Ee+-]+)", r.stdout).group(1))
        print(f"    first dt {first!r} (reference {float(golden['dt_first'])!r})", flush=True)
        assert abs(first - float(golden["dt_first"])) <= 1e-12 * float(golden["dt_first"])
        assert len(dt) == len(golden["dt"])
        err_dt = float(np.abs(dt - golden["dt"]).max() / golden["dt"].max())
        err_cfl = float(np.abs(cfl - golden["cfl"]).max() / golden["cfl"].max())
        print(f"    dt sequence {err_dt:.1e}, CFL*dt {err_cfl:.1e}", flush=True)
        assert err_dt <= 1e-12 and err_cfl <= 1e-12
        for name, key, tkey in (("t1.u", "u_save", "t_save"), ("end.u", "u_end", "t_end")):
            uf = ufile.read_u(os.path.join(work, name))
            assert abs(uf.t - float(golden[tkey])) <= 1e-12 * float(golden[tkey])
            u = np.delete(uf.u, 16, axis=1)
            rr = np.delete(golden[key], 16, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_writes_pressure_files():
    """`blagpu --pressure` writes name.p beside every saved field (wdiscp.f's
    layout, what bla writes with par.f pressure = 1): the t1.u.p and t10.u.p
    of the C1 case equal the reference's (tests/golden/p_32x33x32.npz) to
    1e-9, the pressure bar (its mean mode is an O(1e-5) balance of O(1)
    terms)."""
    _need_binary()
    print("[gpu] blagpu --pressure 32x33x32", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "p_32x33x32.npz"))
    work = _c1_dir(maxit=30, iamp=0)
    try:
        r = subprocess.run([BLAGPU, "--pressure"], cwd=work, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u.p", "p1"), ("t10.u.p", "p10")):
            pf = ufile.read_u(os.path.join(work, name), ncomp=1)
            assert (pf.nx, pf.nyp, pf.nz) == (32, 33, 32)
            err = _rel(np.delete(pf.u[0], 16, axis=0), np.delete(golden[key], 16, axis=0))
            print(f"    {name}: rel L2 {err:.2e}", flush=True)
            assert err <= 1e-9, (name, err)
        assert os.path.exists(os.path.join(work, "end.u.p"))
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
@pytest.mark.parametrize("ibc", [101, 100])
def test_fortran_host_blasius_matches_reference_bla(ibc):
    """The host on the spatial Blasius boundary layer (fltype 6, ibc 101,
    fringe and lower-wall strip: case b of tests/golden/blasius_64x33x8.npz;
    ibc 100 without the strip: tests/golden/blasius100_64x33x8.npz) with
    bla.i in displacement thicknesses, the .u file in the units of dstar =
    2/yl and fsc.dat from the library's similarity solver (equal to the
    reference fsc's): it writes the fields the reference bla wrote."""
    from simson_amd import baseflow
    _need_binary()
    print(f"[gpu] blagpu spatial Blasius boundary layer 64x33x8, ibc {ibc}", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "blasius_64x33x8.npz"))
    g100 = np.load(os.path.join(ROOT, "tests", "golden", "blasius100_64x33x8.npz"))
    ds = 2.0 / 20.0
    work = tempfile.mkdtemp(prefix="blagpu_bl.")
    try:
        baseflow.write_fsc(os.path.join(work, "fsc.dat"), baseflow.fsc_table())
        f = ufile.UField(u=golden["init"], re=400.0 / ds, xl=100.0 * ds, zl=20.0 * ds, t=0.0,
                         xs=0.0, fltype=6, dstar=ds,
                         extra={"bstart": -35.0 * ds, "blength": 25.0 * ds})
        ufile.write_u(os.path.join(work, "init.u"), f)
        wall = dict(wbci=1, wamp=0.01, wdamp=1.0, wxstart=-20.0, wxend=-10.0, wxrise=3.0,
                    wxfall=3.0, wzbet=0.5, wtomeg=0.3) if ibc == 101 else {}
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60, dt=0.1,
                           nst=3, xl=100.0, cflux=False, retau=0.0, ibc=ibc, icfl=3,
                           spat=True, fmax=1.0, fstart=-45.0, fend=0.0, frise=15.0,
                           ffall=10.0, tsave=[0.1], nmsave=["t1.u"], **wall)
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(p))
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        ref = {"t1.u": golden["b_u1"], "end.u": golden["b_u20"]} if ibc == 101 else \
            {"t1.u": g100["u1"], "end.u": g100["u20"]}
        for name in ("t1.u", "end.u"):
            uf = ufile.read_u(os.path.join(work, name))
            assert uf.fltype == 6 and uf.extra["bstart"] == -35.0 * ds
            u = np.delete(uf.u, 4, axis=1)
            rr = np.delete(ref[name], 4, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_temporal_blasius_matches_reference_bla():
    """The host on the temporal Blasius boundary layer (fltype 3, ibc 101,
    the parallel-flow forcing; tests/golden/tblasius_32x33x16.npz) writes
    the fields the reference bla wrote."""
    from simson_amd import baseflow
    _need_binary()
    print("[gpu] blagpu temporal Blasius boundary layer 32x33x16", flush=True)
    golden = np.load(os.path.join(ROOT, "tests", "golden", "tblasius_32x33x16.npz"))
    ds = 2.0 / 20.0
    work = tempfile.mkdtemp(prefix="blagpu_tbl.")
    try:
        baseflow.write_fsc(os.path.join(work, "fsc.dat"), baseflow.fsc_table())
        f = ufile.UField(u=golden["init"], re=400.0 / ds, xl=60.0 * ds, zl=30.0 * ds, t=0.0,
                         xs=0.0, fltype=3, dstar=ds)
        ufile.write_u(os.path.join(work, "init.u"), f)
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60, dt=0.1,
                           nst=3, xl=60.0, cflux=False, retau=0.0, ibc=101, icfl=3,
                           spat=False, cdev=0.0, tsave=[0.1], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(p))
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u", "u1"), ("end.u", "u20")):
            uf = ufile.read_u(os.path.join(work, name))
            assert uf.fltype == 3
            u = np.delete(uf.u, 8, axis=1)
            rr = np.delete(golden[key], 8, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)


@pytest.mark.gpu
def test_fortran_host_local_force_matches_reference_bla():
    """The host with bla.i loctyp = 3 on the spatial Blasius boundary layer
    (tests/golden/locf_64x33x8.npz) writes the fields the reference wrote."""
    from simson_amd import baseflow
    _need_binary()
    print("[gpu] blagpu loctyp 3 volume force 64x33x8", flush=True)
    gb = np.load(os.path.join(ROOT, "tests", "golden", "blasius_64x33x8.npz"))
    golden = np.load(os.path.join(ROOT, "tests", "golden", "locf_64x33x8.npz"))
    ds = 2.0 / 20.0
    work = tempfile.mkdtemp(prefix="blagpu_locf.")
    try:
        baseflow.write_fsc(os.path.join(work, "fsc.dat"), baseflow.fsc_table())
        f = ufile.UField(u=gb["init"], re=400.0 / ds, xl=100.0 * ds, zl=20.0 * ds, t=0.0,
                         xs=0.0, fltype=6, dstar=ds,
                         extra={"bstart": -35.0 * ds, "blength": 25.0 * ds})
        ufile.write_u(os.path.join(work, "init.u"), f)
        p = blai.BlaParams(namnin="init.u", namnut="end.u", tmax=1.0e6, maxit=60, dt=0.1,
                           nst=3, xl=100.0, cflux=False, retau=0.0, ibc=101, icfl=3,
                           spat=True, fmax=1.0, fstart=-45.0, fend=0.0, frise=15.0,
                           ffall=10.0, loctyp=3, lamp=(0.0, 0.01, 0.005), lxscale=8.0,
                           lxloc0=-15.0, lyscale=2.0, lxtype=0.0, ltomeg=0.3,
                           tsave=[0.1], nmsave=["t1.u"])
        with open(os.path.join(work, "bla.i"), "w") as fh:
            fh.write(blai.format_channel(p))
        r = _run(work)
        assert r.returncode == 0, r.stdout + r.stderr
        for name, key in (("t1.u", "u1"), ("end.u", "u20")):
            u = np.delete(ufile.read_u(os.path.join(work, name)).u, 4, axis=1)
            rr = np.delete(golden[key], 4, axis=1)
            for c in range(3):
                err = _rel(u[c], rr[c])
                print(f"    {name} comp {c}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-12, (name, c, err)
    finally:
        shutil.rmtree(work)
