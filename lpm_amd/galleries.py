"""Initial conditions and closed-form solutions on the unit sphere (numpy).

Inputs of the direct sum and the analytic references its results are
reported against.  Points are arrays of shape (3, n) (SoA) or (n, 3) when
``aos=True``.
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2 * np.pi


def _xyz(x):
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 2 and x.shape[0] == 3:
        return x[0], x[1], x[2]
    return x[..., 0], x[..., 1], x[..., 2]


def longitude(x):
    """Longitude in [0, 2 pi) (atan4, src/util/lpm_math.hpp:70-104)."""
    x0, x1, _ = _xyz(x)
    lon = np.arctan2(x1, x0)
    return np.where(lon < 0, lon + TWO_PI, lon)


def latitude(x):
    x0, x1, x2 = _xyz(x)
    return np.arctan2(x2, np.sqrt(x0 * x0 + x1 * x1))


def solid_body_vorticity(x, omega=TWO_PI):
    """SolidBodyRotation: zeta = 2 Omega z (src/lpm_vorticity_gallery.hpp:31-57)."""
    return 2 * omega * _xyz(x)[2]


def solid_body_velocity(x, omega=TWO_PI):
    x0, x1, _ = _xyz(x)
    return np.stack([-omega * x1, omega * x0, np.zeros_like(x0)])


def rh54_vorticity(x, u0, amp):
    """RossbyHaurwitz54: 2 u0 z + 30 amp cos(4 lon) z (z^2-1)^2
    (src/lpm_vorticity_gallery.hpp:104-147).  The Rossby-Haurwitz wave with
    zeta = 2 w sin(lat) - 30 K cos^4(lat) sin(lat) cos(4 lon) is u0 = w,
    amp = -K."""
    z = _xyz(x)[2]
    return 2 * u0 * z + 30 * amp * np.cos(4 * longitude(x)) * z * (z * z - 1) ** 2


def rh54_velocity(x, u0, amp):
    """RossbyWave54Velocity at t = 0 (src/lpm_velocity_gallery.hpp:250-288)."""
    x0, x1, _ = _xyz(x)
    lat = latitude(x)
    lon = longitude(x)
    cl = np.cos(lat)
    u = amp * 0.5 * np.cos(4 * lon) * cl ** 3 * (5 * np.cos(2 * lat) - 3)
    v = amp * 4 * cl ** 3 * np.sin(lat) * np.sin(4 * lon)
    return np.stack([
        -u0 * x1 - u * np.sin(lon) - v * np.sin(lat) * np.cos(lon),
        u0 * x0 + u * np.cos(lon) - v * np.sin(lat) * np.sin(lon),
        v * np.cos(lat)])


def rh54_stream_function(x, u0, amp):
    """psi with Laplacian(psi) = -zeta: u0 z + amp cos(4 lon) z (z^2-1)^2
    (Y_5^4 has eigenvalue -30)."""
    z = _xyz(x)[2]
    return u0 * z + amp * np.cos(4 * longitude(x)) * z * (z * z - 1) ** 2


def rh54_wave_speed(u0, Omega, R=4):
    """Angular phase speed (R(3+R) u0 - 2 Omega) / ((1+R)(2+R)) of the
    Rossby-Haurwitz wave of wavenumber R."""
    return (R * (3 + R) * u0 - 2 * Omega) / ((1 + R) * (2 + R))


def rotate_z(x, angle):
    """Rotate points about the z axis by ``angle`` (SoA (3, n))."""
    c, s = np.cos(angle), np.sin(angle)
    x0, x1, x2 = _xyz(x)
    return np.stack([c * x0 - s * x1, s * x0 + c * x1, x2])


def gaussian_vortex_vorticity(x, ctr, strength=4 * np.pi, b=4.0, gauss_const=0.0):
    """GaussianVortexSphere: strength exp(-b^2 (1 - x.ctr)) - gauss_const
    (src/lpm_vorticity_gallery.hpp:59-102)."""
    x0, x1, x2 = _xyz(x)
    distsq = 1.0 - x0 * ctr[0] - x1 * ctr[1] - x2 * ctr[2]
    return strength * np.exp(-b * b * distsq) - gauss_const


def tc2_state(x, u0=0.2, Omega=TWO_PI, g=1.0, h0=1.0):
    """Steady zonal geostrophic flow u = u0 cos(lat) (Williamson test case 2
    form of examples/sphere_swe_tc2.cpp:205-297): velocity, relative
    vorticity 2 u0 z, zero divergence, surface
    s = h0 + (Omega u0 + u0^2/2) cos^2(lat) / g (h0 = polar depth),
    double dot -2 u0^2 z^2 and surface Laplacian
    (u0^2 + 2 Omega u0)(2 sin^2 - cos^2)/g."""
    x0, x1, x2 = _xyz(x)
    cos2 = 1 - x2 * x2
    sin2 = x2 * x2
    return {
        "u": np.stack([-u0 * x1, u0 * x0, np.zeros_like(x0)]),
        "zeta": 2 * u0 * x2,
        "sigma": np.zeros_like(x0),
        "surface": h0 + (Omega * u0 + 0.5 * u0 * u0) * cos2 / g,
        "ddot": -2 * u0 * u0 * x2 * x2,
        "lap_surface": (u0 * u0 + 2 * Omega * u0) * (2 * sin2 - cos2) / g,
    }


# ------------------------------------------------------------ random inputs
_MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 stream; uniform doubles (r >> 11) * 2^-53 (compiler and
    platform independent)."""

    def __init__(self, seed: int):
        self.state = seed & _MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def uniform(self, n: int = None):
        if n is None:
            return (self.next_u64() >> 11) * (1.0 / (1 << 53))
        return self.uniform_array(n)

    def uniform_array(self, n: int) -> np.ndarray:
        """n uniforms in [0, 1), vectorised (same stream as n scalar draws)."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & _MASK64
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def shoemake_rotation(u1, u2, u3):
    """Rotation matrix of the unit quaternion drawn from three uniforms by
    Shoemake's method (uniform over SO(3)):
    q = (sqrt(1-u1) sin 2pi u2, sqrt(1-u1) cos 2pi u2,
         sqrt(u1) sin 2pi u3, sqrt(u1) cos 2pi u3) = (x, y, z, w)."""
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    qx, qy = a * np.sin(TWO_PI * u2), a * np.cos(TWO_PI * u2)
    qz, qw = b * np.sin(TWO_PI * u3), b * np.cos(TWO_PI * u3)
    return np.array([
        [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
        [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
        [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)],
    ])


ENSEMBLE_SEED = 0x4C504D34


def ensemble_member(m: int, seed_base: int = ENSEMBLE_SEED, levels=(3, 6)):
    """Parameters of ensemble member m: icosahedral level
    L = lo + (r mod (hi - lo + 1)), Rossby-Haurwitz 5-4 with omega and K
    uniform in [0.05, 0.15), and a uniformly random rotation R of the
    pattern; all drawn from splitmix64(seed_base + m) in that order."""
    rng = SplitMix64(seed_base + m)
    lo, hi = levels
    level = lo + rng.next_u64() % (hi - lo + 1)
    omega = 0.05 + 0.1 * rng.uniform()
    K = 0.05 + 0.1 * rng.uniform()
    R = shoemake_rotation(rng.uniform(), rng.uniform(), rng.uniform())
    return {"level": int(level), "omega": omega, "K": K, "rotation": R}


def rotated_rh54_vorticity(x, omega, K, R):
    """Rossby-Haurwitz 5-4 vorticity (u0 = omega, amp = -K) of the pattern
    rotated by R: zeta(x) = zeta_RH(R^T x)."""
    x = np.asarray(x, dtype=np.float64)
    return rh54_vorticity(R.T @ x, omega, -K)
