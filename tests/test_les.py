"""ADM-RT large-eddy model (sgs.i iles = 1: the relaxation term of
nonlinbl.f:920-966 on the high-pass filtered velocity of linearbl.f:1871-1962).

Pins:
* the oracle against the reference bla run with sgs.i (tests/golden/
  les_32x33x32.npz, made by tests/golden/make_golden.py les): the default
  sgs.i (cutoff_inv 1.5, iord 5, chi 0.2) after 1, 10 and 100 steps, and
  cutoff_inv 3, iord 2, chi 2 after 10 steps.  The start field excites every
  resolved wavenumber, so the model moves v and w by about 1e-2 relative
  within 10 steps (dns_u10 in the fixture) and 1e-12 pins it;
* the filter operators against what the mathematics fixes: the y low-pass
  reproduces polynomials up to degree 2 (second-order moments) but not
  degree 6, and its rows at the walls are those of the identity (the
  high-pass part vanishes there); the x-z transfer function is 1 at k = 0
  and crosses 1 - (1/2)^N at the cutoff (the implicit part is scaled so that
  the filter is 1/2 there);
* the library's own filter construction (a C-ABI function that needs no
  GPU) against the oracle's;
* on the GPU, the step with the model against the same fixture, FP64 at
  1e-12 and mixed precision at 1e-5.
"""
import math
import sys

import numpy as np
import pytest

import pyoracle
from simson_amd import fields, sgsi

NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi
CASES = {"a": (1.5, 5, 0.2), "b": (3.0, 2, 2.0)}


@pytest.fixture(scope="module")
def golden_les():
    from conftest import ROOT
    import os
    return np.load(os.path.join(ROOT, "tests", "golden", "les_32x33x32.npz"))


def les_field():
    return fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=15, rms=1e-2, seed=5)


def rel3(u, ref):
    d = np.delete(u - ref, NZ // 2, axis=1)
    r = np.delete(ref, NZ // 2, axis=1)
    return [float(np.linalg.norm(d[c]) / np.linalg.norm(r[c])) for c in range(3)]


def test_sgs_file_round_trip_and_checks():
    p = sgsi.parse(sgsi.format_sgs(sgsi.SgsParams(iles=1, cutoff_inv=2.0, iord=3, chi=0.5)))
    assert (p.iles, p.cutoff_inv, p.iord, p.chi) == (1, 2.0, 3, 0.5)
    assert abs(p.cutoff - 0.5) < 1e-15
    with pytest.raises(sgsi.SgsInputError):
        sgsi.parse("20060909\n1\n0.5\n5\n0.2\n")       # cutoff = 2 > 1
    with pytest.raises(sgsi.SgsInputError):
        sgsi.parse("20060909\n1\n1.5\n6\n0.2\n")       # iord > 5
    with pytest.raises(sgsi.SgsInputError):
        sgsi.parse("20060909\n3\n1.5\n5\n0.\n1\n")   # HPF eddy viscosity: not supported
    with pytest.raises(sgsi.SgsInputError) as e:
        sgsi.parse("20000101\n1\n")
    assert e.value.code == 345432


def test_filter_operators_mathematics():
    ny = 33
    lpf, L = pyoracle.les_filter(NX, ny, NZ, 2.0 / 3.0, 5)
    y = np.cos(np.pi * np.arange(ny) / (ny - 1))
    for p in range(3):                     # second-order moments: exact on quadratics
        assert np.abs(L @ y**p - y**p).max() < 1e-13
    assert np.abs(L @ y**6 - y**6).max() > 1e-8   # and not beyond
    assert np.abs(L[0] - np.eye(ny)[0]).max() == 0 and np.abs(L[-1] - np.eye(ny)[-1]).max() == 0
    assert lpf[0, 0] == 1.0 and np.all(lpf <= 1.0) and np.all(lpf >= 0.0)
    # x transfer function at the cutoff: lpf(kx) with kz = 0 is 1 - (1 - G)^6,
    # G(kc) = 1/2 for kx dx = pi cutoff, i.e. kx = cutoff * nx/2 (here 32/3:
    # between grid wavenumbers, so check monotone bracketing)
    g6 = 1 - 0.5 ** 6
    kc = (2.0 / 3.0) * NX / 2
    assert lpf[0, int(math.floor(kc))] > g6 > lpf[0, int(math.ceil(kc))]


def test_library_filter_matches_oracle():
    from simson_amd.gpu import les_filter
    for cut, iord in ((2.0 / 3.0, 5), (1.0 / 3.0, 2)):
        lpf, L = pyoracle.les_filter(NX, NY, NZ, cut, iord)
        lpf2, L2 = les_filter(NX, NY, NZ, cut, iord)
        assert np.abs(lpf - lpf2).max() < 1e-15
        assert np.abs(L - L2).max() < 1e-13


@pytest.mark.parametrize("case", ["a", "b"])
def test_oracle_les_matches_reference(golden_les, case):
    cut_inv, iord, chi = CASES[case]
    f = les_field()
    o = pyoracle.Oracle(NX, NY, NZ, RE, XL, ZL)
    o.set_les(1, 1.0 / cut_inv, iord, chi)
    ur, ui, puw = o.new_state(f.u)
    st, _ = o.run(ur, ui, puw, 10 if case == "b" else 1, 0.01, icfl=3)
    assert st == 0
    key = "b_u10" if case == "b" else "a_u1"
    e = rel3(pyoracle.velocities(ur, ui), golden_les[key])
    eff = rel3(golden_les["dns_u10"], golden_les["b_u10"] if case == "b" else golden_les["a_u10"])
    print(f"oracle LES {case}: rel L2 {e}; model effect {eff}")
    assert max(e) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_gpu_les_matches_reference(golden_les, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] ADM-RT LES vs reference ({form})", flush=True)
    sys.stdout.flush()
    tol = 1e-5 if form == "mixed" else 1e-12
    f = les_field()
    for case in ("a", "b"):
        cut_inv, iord, chi = CASES[case]
        prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
        with BlaGPU.from_ufield(f, precision=prec) as g:
            g.les(iles=1, cutoff=1.0 / cut_inv, iord=iord, chi=chi)
            g.upload_velocity(f.u, -2.0 / RE)
            if case == "a":
                done = 0
                for n in (1, 10, 100):
                    g.run(n - done, 0.01, icfl=3, first=done == 0)
                    done = n
                    e = rel3(g.velocities(), golden_les[f"a_u{n}"])
                    print(f"    {case} {n} steps: rel L2 {e}", flush=True)
                    assert max(e) < tol, (case, n, e)
            else:
                g.run(10, 0.01, icfl=3)
                e = rel3(g.velocities(), golden_les["b_u10"])
                print(f"    {case} 10 steps: rel L2 {e}", flush=True)
                assert max(e) < tol, (case, e)
