"""ADM-RT large-eddy model (sgs.i iles = 1: the relaxation term of
nonlinbl.f:920-966 on the high-pass filtered velocity of linearbl.f:1871-1962).

Pins:
* the oracle against the reference bla run with sgs.i (tests/golden/
  les_32x33x32.npz, made by tests/golden/make_golden.py les): the default
  sgs.i (cutoff_inv 1.5, iord 5, chi 0.2) after 1, 10 and 100 steps, and
  cutoff_inv 3, iord 2, chi 2 after 10 steps.  The start field excites every
  resolved wavenumber, so the model moves v and w by about 1e-2 relative
  within 10 steps (dns_u10 in the fixture) and 1e-12 pins it;
* the row stencils (oracle/les_ref.py, derived from their conditions)
  against those conditions checked one by one: moments 0-2, B(pi) = 0, the
  third moment (wall rows), minimum dispersion by brute force over the
  one-parameter family the other conditions leave (interior rows), and the
  implicit factor's |G(pi cutoff)| = 1/2;
* the filter operators against what the mathematics fixes: the y low-pass
  reproduces polynomials up to degree 2 (second-order moments) but not
  degree 6, and its rows at the walls are those of the identity (the
  high-pass part vanishes there); the x-z transfer function is 1 at k = 0
  and crosses 1 - (1/2)^N at the cutoff (the implicit part is scaled so that
  the filter is 1/2 there);
* the library's own filter construction (a C-ABI function that needs no
  GPU; a Lagrange system and dense pivoted solves in C++) against the
  oracle's (SVD null space and numpy solves);
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
        sgsi.parse("20060909\n4\n")                     # no such model
    with pytest.raises(sgsi.SgsInputError) as e:
        sgsi.parse("20000101\n1\n")
    assert e.value.code == 345432


def test_row_stencils_satisfy_their_conditions():
    import les_ref
    ny = 33
    eta = np.cos(np.pi * np.arange(ny) / (ny - 1))
    for cutoff in (2.0 / 3.0, 1.0 / 3.0, 0.9):
        for k in (1, 2, 5, ny // 2, ny - 3, ny - 2):
            cols = (np.arange(0, 5) if k == 1 else np.arange(ny - 5, ny) if k == ny - 2
                    else np.arange(k - 2, k + 3))
            j = cols - k
            p = eta[k] - eta[cols]
            wall = k in (1, ny - 2)
            b, a = les_ref.stencil(p, j, wall, cutoff)
            h = np.abs(p).max()
            mom = [float(b @ (p / h) ** m) for m in range(4)]
            assert abs(mom[0] - 1) < 1e-14 and abs(mom[1]) < 1e-14 and abs(mom[2]) < 1e-14
            assert abs(b @ (-1.0) ** j) < 1e-14                  # B(pi) = 0
            if wall:
                assert abs(mom[3]) < 1e-13
            else:
                # brute force over the family b + t n (n: null vector of the
                # four other conditions): the odd part is smallest at t = 0
                Cm = np.array([np.ones(5), p / h, (p / h) ** 2, (-1.0) ** j])
                n = np.linalg.svd(Cm)[2][-1]
                odd = lambda w: (w[3] - w[1]) ** 2 + (w[4] - w[0]) ** 2
                ts = np.linspace(-1, 1, 2001) * 0.05
                best = ts[np.argmin([odd(b + t * n) for t in ts])]
                assert abs(best) <= 0.05 / 1000 + 1e-15, (k, cutoff, best)
            # implicit part: a_j = r b_j off the centre, weights sum to 1,
            # and |B/A| = 1/2 at the cutoff
            c = list(j).index(0)
            r = a[(c + 1) % 5] / b[(c + 1) % 5]
            assert np.allclose(np.delete(a, c), r * np.delete(b, c), rtol=1e-13, atol=0)
            assert abs(a.sum() - 1) < 1e-14
            kc = np.pi * cutoff
            G = les_ref.transfer(b, j, kc)[0] / les_ref.transfer(a, j, kc)[0]
            assert abs(abs(G) - 0.5) < 1e-13, (k, cutoff, abs(G))
    # the homogeneous filter is real, 1 at k = 0, 0 at the grid Nyquist
    g = les_ref.homogeneous_transfer(17, 2.0 / 3.0)
    assert abs(g[0] - 1) < 1e-15 and abs(g[-1]) < 1e-15
    with pytest.raises(ValueError):
        les_ref.filter_operators(NX, NY, NZ, 1.0, 5)


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
    for ny, cut, iord in ((NY, 2.0 / 3.0, 5), (NY, 1.0 / 3.0, 2), (129, 2.0 / 3.0, 5)):
        lpf, L = pyoracle.les_filter(NX, ny, NZ, cut, iord)
        lpf2, L2 = les_filter(NX, ny, NZ, cut, iord)
        assert np.abs(lpf - lpf2).max() < 1e-14
        assert np.abs(L - L2).max() < 1e-13
    from simson_amd.gpu import SimsonError
    with pytest.raises(SimsonError):
        les_filter(NX, NY, NZ, 1.0, 5)


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
