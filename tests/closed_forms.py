"""Closed-form vorticity of single tubes whose deposition the mathematics fixes.

Each case returns run parameters, the tube, and the field the reference's deposition formula
(tube.f90:844-892) gives for that geometry, written from the geometry itself (a straight line or a
circle), not from the deposition code:

  straight branch (|c' x c''| < 1e-8, tube.f90:854-860):
      omega = Gamma tp [T + eta (xc x T) + (rho dsigma/ds / sigma) e_rho]
  curved branch (tube.f90:868-892), Frenet frame N = B x T, D = 1 - kappa rho cos(theta):
      omega = Gamma tp [T - eta (xc x T) / D + (dsigma/ds / (sigma D)) xc]

with tp = [exp(-rho^2 / 2 sigma^2) - exp(-R^2 / 2 sigma^2)] / (2 pi sigma^2), xc the offset of the
cell from the centerline point in its normal plane, rho = |xc|. For a circle of radius a with N
pointing to the centre, D = r / a (r: the cell's distance from the circle's axis). The two
branches turn the twist term in opposite senses: the straight branch adds eta rho (e_rho x T)
(tube.f90:857-859 with e_th = e_rho x T), the curved one eta (t5 N + t6 B) = -eta rho (e_rho x T) / D
(tube.f90:878-880); both are the reference's statements as written.

sigma(zeta) = smin + (smax - smin)(sin(w zeta) + 1)/2, w = floor(length/pi + 1/2) n_T
(subroutine.f90:164-173), and ds = Len_Disc dzeta along the arc-length spline, so
dsigma/ds = sigma'(zeta) / Len_Disc (tube.f90:851-853). The cut-off R = Rtube_ratio
max(sigma(s0), sigma(s1)) of the sub-segment holding the cell (tube.f90:786-792).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from fastweavturb_amd.woven_io import RunParams, Tubes


@dataclasses.dataclass
class Case:
    name: str
    params: RunParams
    tubes: Tubes
    expect: np.ndarray      # (3, N, N, N) [c, z, y, x]
    rtol_l2: float          # relative L2 bound of (field - expect)
    atol_max: float         # bound of max |field - expect| relative to max |expect|
    make: object = None     # the factory and its arguments, to rebuild the case with a mistake
    kw: dict = dataclasses.field(default_factory=dict)

    def with_mistake(self, mistake: str) -> "Case":
        """The same case with one plausible mistake planted in the expected field: "flip_twist"
        (twist term of the opposite sign), "drop_core" (no varying-core term), "drop_D" (curved
        branch without the 1/D of the Frenet metric)."""
        return self.make(**self.kw, mistake=mistake)


def grid(N):
    x = -np.pi + 2 * np.pi * np.arange(N) / N
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def _minimg(d):
    return np.where(d > np.pi, d - 2 * np.pi, np.where(d < -np.pi, d + 2 * np.pi, d))


def _sigma(zeta, length, smin, ratio, n_T):
    w = np.floor(length / np.pi + 0.5) * n_T
    smax = ratio * smin
    return smin + (smax - smin) * (np.sin(w * zeta) + 1) / 2, (smax - smin) / 2 * w * np.cos(w * zeta)


def _tp(rho, sig, R):
    return np.where(rho < R, (np.exp(-rho ** 2 / (2 * sig ** 2)) - np.exp(-R ** 2 / (2 * sig ** 2)))
                    / (2 * np.pi * sig ** 2), 0.0)


def _cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def straight(N=32, npts=40, sigma=0.15, gamma=1.3, eta=0.0, ratio=1.0, n_T=1, y0=0.1234, z0=-0.4321,
             name="straight", mistake=None) -> Case:
    """Periodic straight tube along +x through (y0, z0): points x_i = -pi + (i + 0.37) h, h = 2 pi/npts
    (no point on a grid plane, so every cell lies strictly inside one section's slab). The centerline
    has no turning, so ntf = 1 and the sub-segments are the sections; the spline is
    x(zeta) = x_0 + zeta exactly, Len_Disc = 1, length = 2 pi."""
    h = 2 * np.pi / npts
    x = -np.pi + (np.arange(npts) + 0.37) * h
    t = Tubes(np.array([npts], np.int32), x, np.full(npts, y0), np.full(npts, z0), np.array([sigma]),
              np.array([gamma]), np.array([n_T], np.int32), np.array([0], np.int32))
    p = RunParams(N, N, 1, 1, ratio, eta, 50, 200, 3.0, 1e-4)
    X, Y, Z = grid(N)
    zeta = np.mod(X - x[0], 2 * np.pi)
    sec = np.minimum(np.floor(zeta / h).astype(int), npts - 1)
    sig, dsig = _sigma(zeta, 2 * np.pi, sigma, ratio, n_T)
    sa, _ = _sigma(sec * h, 2 * np.pi, sigma, ratio, n_T)
    sb, _ = _sigma((sec + 1) * h, 2 * np.pi, sigma, ratio, n_T)
    R = 3.0 * np.maximum(sa, sb)
    xc = np.stack([np.zeros_like(X), _minimg(Y - y0), _minimg(Z - z0)])
    rho = np.sqrt(xc[1] ** 2 + xc[2] ** 2)
    T = np.stack([np.ones_like(X), np.zeros_like(X), np.zeros_like(X)])
    tp = _tp(rho, sig, R)
    tw = -eta if mistake == "flip_twist" else eta
    cf = 0.0 if mistake == "drop_core" else 1.0
    om = gamma * tp * (T + tw * _cross(xc, T) + cf * (dsig / sig) * xc)
    kw = dict(N=N, npts=npts, sigma=sigma, gamma=gamma, eta=eta, ratio=ratio, n_T=n_T, y0=y0, z0=z0, name=name)
    # the reference's central difference for dsigma/dzeta (step 1e-4) is within (w 1e-4)^2/6 relative
    return Case(name, p, t, om, 5e-8, 5e-8, straight, kw)


def ring(N=32, npts=256, a=1.5, centre=(0.1, -0.2, 0.3), sigma=0.25, gamma=1.1, eta=0.0, ratio=1.0, n_T=1,
         ntf_times=8000, name="ring", mistake=None) -> Case:
    """Circle of radius a in the plane z = centre[2], traversed counter-clockwise from the angle 0,
    npts uniform points. The chords are equal, so zeta is the angle; Len_Disc = a (to the chord
    polygon's length); the Frenet frame is T = e_phi, N = -e_r, B = +e_z, kappa = 1/a.
    The reference interpolates zeta linearly along each chord rather than projecting on the
    circle, which moves the centerline point by at most (R/a) (dphi/2) a along the tangent
    (dphi: a sub-segment's angle); ntf_times = 8000 makes dphi = 2 pi/npts/33."""
    cx0, cy0, z0 = centre
    phi = 2 * np.pi * np.arange(npts) / npts
    t = Tubes(np.array([npts], np.int32), cx0 + a * np.cos(phi), cy0 + a * np.sin(phi), np.full(npts, z0),
              np.array([sigma]), np.array([gamma]), np.array([n_T], np.int32), np.array([0], np.int32))
    p = RunParams(N, N, 1, 1, ratio, eta, ntf_times, 400, 3.0, 1e-4)
    X, Y, Z = grid(N)
    dX, dY = X - cx0, Y - cy0
    r = np.hypot(dX, dY)
    ph = np.arctan2(dY, dX)
    T = np.stack([-np.sin(ph), np.cos(ph), np.zeros_like(X)])
    foot = np.stack([cx0 + a * np.cos(ph), cy0 + a * np.sin(ph), np.full_like(X, z0)])
    xc = np.stack([X, Y, Z]) - foot
    rho = np.sqrt((xc ** 2).sum(axis=0))
    length = npts * 2 * a * np.sin(np.pi / npts)          # chord polygon
    LD = length / (2 * np.pi)
    zeta = np.mod(ph, 2 * np.pi)
    sig, dsig = _sigma(zeta, length, sigma, ratio, n_T)
    ntf = int(np.ceil(ntf_times / npts)) + 1
    dz = 2 * np.pi / npts / ntf
    j = np.floor(zeta / dz)
    sa, _ = _sigma(j * dz, length, sigma, ratio, n_T)
    sb, _ = _sigma((j + 1) * dz, length, sigma, ratio, n_T)
    R = 3.0 * np.maximum(sa, sb)
    D = np.ones_like(r) if mistake == "drop_D" else r / a
    dsds = dsig / LD
    tp = _tp(rho, sig, R)
    tw = -eta if mistake == "flip_twist" else eta
    cf = 0.0 if mistake == "drop_core" else 1.0
    om = gamma * tp * (T - tw * _cross(xc, T) / D + cf * (dsds / (sig * D)) * xc)
    kw = dict(N=N, npts=npts, a=a, centre=centre, sigma=sigma, gamma=gamma, eta=eta, ratio=ratio, n_T=n_T,
              ntf_times=ntf_times, name=name)
    return Case(name, p, t, om, 1e-3, 2e-3, ring, kw)


def cases() -> list[Case]:
    return [
        straight(eta=1.7, name="straight_twist"),
        straight(ratio=4.0, n_T=2, name="straight_varying_core"),
        straight(ratio=4.0, n_T=1, eta=-0.9, sigma=0.12, name="straight_twist_varying_core"),
        ring(name="ring"),
        ring(eta=1.5, name="ring_twist"),
        ring(eta=1.5, ratio=3.0, n_T=2, sigma=0.16, name="ring_twist_varying_core"),
    ]


def check(case: Case, om: np.ndarray) -> tuple[float, float]:
    """Relative L2 and relative max error of om against the closed form; raises if above bounds."""
    d = om - case.expect
    l2 = float(np.linalg.norm(d) / np.linalg.norm(case.expect))
    mx = float(np.abs(d).max() / np.abs(case.expect).max())
    assert l2 <= case.rtol_l2 and mx <= case.atol_max, (case.name, l2, mx)
    return l2, mx
