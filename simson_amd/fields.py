"""Seeded divergence-free initial fields for the temporal channel.

The generator works mode by mode in (kx, kz) with analytic y polynomials,
so the field satisfies no-slip and continuity exactly (up to rounding):

* for every stored mode with 0 < |kx|,|kz| <= kmax (kx >= 0 in storage),
  v(y) = (1-y^2)^2 (a0 + a1 y + a2 y^2) and omega_y(y) = (1-y^2)(b0 + b1 y)
  with complex coefficients drawn uniformly from [-1, 1];
* u and w follow from continuity and the definition of omega_y,
  u = (i alpha v' - i beta omega_y)/k^2, w = (i beta v' + i alpha omega_y)/k^2
  (the same relations the solver uses, ``bla/evald.f:119-121``);
* the kx = 0 plane is Hermitian: the kz < 0 entry is the conjugate of kz > 0;
* the mean (0,0) mode carries the laminar Poiseuille profile U = 1 - y^2.

The perturbation is scaled to the requested rms, defined as
sqrt((<u'^2> + <v'^2> + <w'^2>)/3) with <.> the volume average.

Random numbers come from splitmix64 so the field is identical on any host.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import polynomial as P

from .ufile import UField

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [-1, 1)."""
        return (self.next_u64() >> 11) * (2.0 / (1 << 53)) - 1.0

    def cuniform(self) -> complex:
        re = self.uniform()
        return complex(re, self.uniform())


def gl_points(nyp: int) -> np.ndarray:
    """Gauss-Lobatto points y_j = cos(pi j/(nyp-1)) (``bla/preprbl.f:44-46``)."""
    return np.cos(np.pi * np.arange(nyp) / (nyp - 1))


def kz_of_index(nz: int) -> np.ndarray:
    iz = np.arange(nz)
    return np.where(iz < nz // 2, iz, iz - nz)


def channel_field(nx: int, nyp: int, nz: int, xl: float, zl: float, re: float,
                  kmax: int, rms: float, seed: int,
                  laminar: bool = True, walls: int = 2) -> UField:
    """Build the seeded perturbed Poiseuille field as a ``UField``.  With
    walls = 1 the envelopes vanish at the lower wall only, (1+y)^2 for v
    and (1+y) for omega_y (a boundary layer with a free stream above), and
    the mean mode is left zero."""
    rng = SplitMix64(seed)
    y = gl_points(nyp)
    if walls == 2:
        env2 = P.polypow([1.0, 0.0, -1.0], 2)      # (1-y^2)^2
        env1 = np.array([1.0, 0.0, -1.0])          # (1-y^2)
    elif walls == 1:
        env2 = P.polypow([1.0, 1.0], 2)            # (1+y)^2
        env1 = np.array([1.0, 1.0])                # (1+y)
        laminar = False
    else:
        raise ValueError("walls must be 1 or 2")
    u = np.zeros((3, nz, nyp, nx // 2), dtype=np.complex128)
    kzs = kz_of_index(nz)
    # Gauss-Legendre rule for the y average of |.|^2 (exact for these degrees)
    gx, gw = np.polynomial.legendre.leggauss(32)
    energy = 0.0
    modes = []
    for ix in range(min(kmax, nx // 2 - 1) + 1):
        for iz in range(nz):
            kz = int(kzs[iz])
            if abs(kz) > kmax or kz == -(nz // 2):
                continue
            if ix == 0 and kz <= 0:
                continue
            a = [rng.cuniform() for _ in range(3)]
            b = [rng.cuniform() for _ in range(2)]
            alfa = 2 * np.pi * ix / xl
            beta = 2 * np.pi * kz / zl
            k2 = alfa * alfa + beta * beta
            vpoly = P.polymul(env2, np.array(a, dtype=complex))
            dvpoly = P.polyder(vpoly)
            opoly = P.polymul(env1, np.array(b, dtype=complex))
            upoly = P.polysub(P.polymul([1j * alfa], dvpoly),
                              P.polymul([1j * beta], opoly)) / k2
            wpoly = P.polyadd(P.polymul([1j * beta], dvpoly),
                              P.polymul([1j * alfa], opoly)) / k2
            # kx > 0: the unstored conjugate at -kx; kx = 0: the stored
            # conjugate at -kz. Either way the mode appears twice.
            mult = 2.0
            for poly in (upoly, vpoly, wpoly):
                vals = P.polyval(gx, poly)
                energy += mult * 0.5 * np.sum(gw * np.abs(vals) ** 2)
            modes.append((ix, iz, upoly, vpoly, wpoly))
    scale = rms / np.sqrt(energy / 3.0) if energy > 0 else 0.0
    for ix, iz, upoly, vpoly, wpoly in modes:
        for c, poly in enumerate((upoly, vpoly, wpoly)):
            vals = scale * P.polyval(y, poly)
            u[c, iz, :, ix] = vals
            if ix == 0:
                u[c, (-kzs[iz]) % nz, :, 0] = np.conj(vals)
    if laminar:
        u[0, 0, :, 0] = 1.0 - y * y
    return UField(u=u, re=re, xl=xl, zl=zl, t=0.0, xs=0.0, fltype=1, dstar=1.0)


def boundary_layer_field(nx: int, nyp: int, nz: int, xl: float, zl: float, re: float,
                         dstar: float, base: np.ndarray, bstart: float, blength: float,
                         kmax: int, rms: float, seed: int, fltype: int = 6) -> UField:
    """A spatial boundary layer for a .u file: the base flow base [3][ny]
    [3nx/2] (physical points x_j = -xl/2 + j xl/nxp, e.g. from
    simson_amd.baseflow.spatial_base_flow) as x Fourier modes kx < nx/2,
    plus the seeded one-wall perturbation of channel_field(walls=1).
    Lengths, re and the blending zone in the units of the .u file."""
    f = channel_field(nx, nyp, nz, xl, zl, re, kmax=kmax, rms=rms, seed=seed, walls=1)
    nxp = 3 * nx // 2
    spec = np.fft.fft(np.asarray(base, np.float64), axis=2)[:, :, :nx // 2] / nxp
    f.u[:, 0, :, :] += spec
    f.fltype = fltype
    f.dstar = dstar
    f.extra = {"bstart": bstart, "blength": blength}
    return f
