"""Seeded synthetic point clouds and source terms for the PFEM workloads.

One portable generator, SplitMix64 mapped to doubles in [0, 1) as
(z >> 11) * 2^-53, feeds every workload so that the GPU path, the CPU
oracle and any other consumer see the same points.  Clouds are returned in
input order as (coords [N, dim] float64, alpha [N] float64).

Workloads (all FP64):
  * ``jittered_square``   : 128 x 128 cell-centred grid on the unit square,
                            h = 1/128, jitter U(-0.3h, 0.3h) per coordinate,
                            alpha = 1.3h  -> N = 16,384
  * ``dam_break``         : fluid column [0, 0.5] x [0, 1] as a 704 x 1407
                            cell-centred grid of spacing h = 1/1407 with
                            jitter +-0.3h, plus 9,472 points uniform in
                            [0.5, 1] x [0, 0.3] (free-surface fragments),
                            alpha = 1.2h  -> N = 1,000,000
  * ``uniform_glibc``     : lives in the tests (needs the C library's rand)
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive SplitMix64 outputs from state = seed."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def jittered_grid(nx: int, ny: int, h: float, jitter: float, seed: int, x0=0.0, y0=0.0) -> np.ndarray:
    """Cell-centred nx x ny grid, index k = j*nx + i, each coordinate moved by
    U(-jitter*h, jitter*h)."""
    i = np.tile(np.arange(nx, dtype=np.float64), ny)
    j = np.repeat(np.arange(ny, dtype=np.float64), nx)
    u = uniform(seed, 2 * nx * ny).reshape(-1, 2)
    x = x0 + (i + 0.5) * h + (2.0 * u[:, 0] - 1.0) * jitter * h
    y = y0 + (j + 0.5) * h + (2.0 * u[:, 1] - 1.0) * jitter * h
    return np.stack([x, y], axis=1)


def jittered_square(n_side: int = 128, seed: int = 0):
    """Unit-square jittered grid, alpha = 1.3h."""
    h = 1.0 / n_side
    pts = jittered_grid(n_side, n_side, h, 0.3, seed)
    return pts, np.full(pts.shape[0], 1.3 * h)


DAM_NX, DAM_NY, DAM_SCATTER = 704, 1407, 9472


def dam_break(nx: int = DAM_NX, ny: int = DAM_NY, n_scatter: int = DAM_SCATTER, seed: int = 1):
    """Dam-break-like cloud: jittered fluid column plus scattered fragments,
    alpha = 1.2h.  Defaults give exactly 1,000,000 points."""
    h = 1.0 / ny
    col = jittered_grid(nx, ny, h, 0.3, seed)
    u = uniform(seed + 0x5EED, 2 * n_scatter).reshape(-1, 2)
    frag = np.stack([0.5 + 0.5 * u[:, 0], 0.3 * u[:, 1]], axis=1)
    pts = np.concatenate([col, frag], axis=0)
    return pts, np.full(pts.shape[0], 1.2 * h)


def dam_break_scaled(n_target: int, seed: int = 1):
    """Smaller dam-break cloud with the same aspect and scatter fraction."""
    s = np.sqrt(n_target / 1_000_000.0)
    nx = max(4, int(round(DAM_NX * s)))
    ny = max(8, int(round(DAM_NY * s)))
    ns = max(1, int(round(DAM_SCATTER * s * s)))
    h = 1.0 / ny
    col = jittered_grid(nx, ny, h, 0.3, seed)
    u = uniform(seed + 0x5EED, 2 * ns).reshape(-1, 2)
    frag = np.stack([0.5 + 0.5 * u[:, 0], 0.3 * u[:, 1]], axis=1)
    pts = np.concatenate([col, frag], axis=0)
    return pts, np.full(pts.shape[0], 1.2 * h)


def fourier_modes(seed: int = 11, n_modes: int = 8):
    """Random modes u = sum_m a_m sin(k_m . x + phi_m), |a_m| = 1, integer
    wave numbers in [1, 4] (times pi) per axis."""
    u = uniform(seed, 4 * n_modes).reshape(n_modes, 4)
    kx = np.pi * (1.0 + np.floor(4.0 * u[:, 0]))
    ky = np.pi * (1.0 + np.floor(4.0 * u[:, 1]))
    phi = 2.0 * np.pi * u[:, 2]
    th = 2.0 * np.pi * u[:, 3]
    return np.stack([kx, ky, phi, np.cos(th), np.sin(th)], axis=1)


def fourier_velocity(xy: np.ndarray, modes: np.ndarray) -> np.ndarray:
    """u = sum_m a_m sin(k_m . x + phi_m) at the points, shape (n, 2): the
    velocity whose divergence divergence_source gives."""
    u = np.zeros((xy.shape[0], 2))
    for kx, ky, phi, ax, ay in modes:
        s = np.sin(kx * xy[:, 0] + ky * xy[:, 1] + phi)
        u[:, 0] += ax * s
        u[:, 1] += ay * s
    return u


def divergence_source(xy: np.ndarray, modes: np.ndarray) -> np.ndarray:
    """f = div u = sum_m (a_m . k_m) cos(k_m . x + phi_m) at the points."""
    f = np.zeros(xy.shape[0])
    for kx, ky, phi, ax, ay in modes:
        f += (ax * kx + ay * ky) * np.cos(kx * xy[:, 0] + ky * xy[:, 1] + phi)
    return f


def jittered_grid3(nx: int, ny: int, nz: int, h: float, jitter: float, seed: int) -> np.ndarray:
    """Cell-centred nx x ny x nz grid, index k = (kz*ny + ky)*nx + kx, each
    coordinate moved by U(-jitter*h, jitter*h)."""
    kx = np.tile(np.arange(nx, dtype=np.float64), ny * nz)
    ky = np.tile(np.repeat(np.arange(ny, dtype=np.float64), nx), nz)
    kz = np.repeat(np.arange(nz, dtype=np.float64), nx * ny)
    u = uniform(seed, 3 * nx * ny * nz).reshape(-1, 3)
    return np.stack([(kx + 0.5) * h + (2.0 * u[:, 0] - 1.0) * jitter * h,
                     (ky + 0.5) * h + (2.0 * u[:, 1] - 1.0) * jitter * h,
                     (kz + 0.5) * h + (2.0 * u[:, 2] - 1.0) * jitter * h], axis=1)


TANK_NX, TANK_NY, TANK_NZ = 215, 108, 86


def sloshing_tank(nx: int = TANK_NX, ny: int = TANK_NY, nz: int = TANK_NZ, seed: int = 2):
    """3D sloshing-tank cloud: a jittered cubic grid (+-0.25h) filling
    [0,1] x [0,0.5] x [0,0.4] (h = 1/nx), then z += 0.05 sin(2 pi x) z/0.4,
    which moves the top surface by 0.05 sin(2 pi x) and keeps the bottom;
    alpha = 1.3h.  Defaults give 1,996,920 points."""
    h = 1.0 / nx
    p = jittered_grid3(nx, ny, nz, h, 0.25, seed)
    p[:, 2] += 0.05 * np.sin(2.0 * np.pi * p[:, 0]) * (p[:, 2] / 0.4)
    return p, np.full(p.shape[0], 1.3 * h)


def sloshing_tank_scaled(n_target: int, seed: int = 2):
    """Smaller tank cloud with the same aspect ratio and spacing rule."""
    s = (n_target / float(TANK_NX * TANK_NY * TANK_NZ)) ** (1.0 / 3.0)
    return sloshing_tank(max(4, int(round(TANK_NX * s))), max(3, int(round(TANK_NY * s))),
                         max(3, int(round(TANK_NZ * s))), seed)


ADVECT_SEED, ADVECT_MODES, ADVECT_DT, ADVECT_STEPS = 33, 16, 1e-3, 200


def stream_modes(seed: int = ADVECT_SEED, n_modes: int = ADVECT_MODES, speed: float = 0.25):
    """Stream function psi = sum_m a_m sin(k_m . x + phi_m) for the advection
    sequence: wave numbers pi*(1 + floor(4u)) per axis, phases 2*pi*u and
    a_m = speed / (n_modes |k_m|), so each mode's velocity amplitude is
    speed / n_modes and |u| <= speed.  Rows (kx, ky, phi, a)."""
    u = uniform(seed, 3 * n_modes).reshape(n_modes, 3)
    kx = np.pi * (1.0 + np.floor(4.0 * u[:, 0]))
    ky = np.pi * (1.0 + np.floor(4.0 * u[:, 1]))
    phi = 2.0 * np.pi * u[:, 2]
    a = speed / (n_modes * np.hypot(kx, ky))
    return np.stack([kx, ky, phi, a], axis=1)


def advect(xy: np.ndarray, modes: np.ndarray, dt: float, steps: int = 1) -> np.ndarray:
    """Forward Euler in u = (dpsi/dy, -dpsi/dx) (host restatement of the
    device advection, for tests)."""
    p = np.array(xy, np.float64, copy=True)
    for _ in range(steps):
        ux = np.zeros(p.shape[0])
        uy = np.zeros(p.shape[0])
        for kx, ky, phi, a in modes:
            c = a * np.cos(kx * p[:, 0] + ky * p[:, 1] + phi)
            ux += ky * c
            uy -= kx * c
        p[:, 0] += dt * ux
        p[:, 1] += dt * uy
    return p


TANK8M = (342, 171, 137)  # 8,012,034 points: the tank at 4x the point count
ADVECT3_SEED = 44


def potential_modes(seed: int = ADVECT3_SEED, n_modes: int = ADVECT_MODES, speed: float = 0.25):
    """Divergence-free 3D field u(x) = sum_m cos(k_m . x + phi_m) b_m, the
    curl of the vector potential sum_m a_m sin(k_m . x + phi_m) (b_m =
    k_m x a_m, so div u = 0 exactly): wave numbers pi*(1 + floor(4u)) per
    axis, phases 2*pi*u, directions a_m uniform on the sphere, scaled so
    |b_m| = speed / n_modes (|u| <= speed).  Rows (kx, ky, kz, phi, bx, by, bz)."""
    u = uniform(seed, 6 * n_modes).reshape(n_modes, 6)
    k = np.pi * (1.0 + np.floor(4.0 * u[:, :3]))
    phi = 2.0 * np.pi * u[:, 3]
    ct = 2.0 * u[:, 4] - 1.0
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    az = 2.0 * np.pi * u[:, 5]
    a = np.stack([st * np.cos(az), st * np.sin(az), ct], axis=1)
    b = np.cross(k, a)
    nb = np.linalg.norm(b, axis=1)
    nb[nb == 0] = 1.0
    b *= (speed / n_modes) / nb[:, None]
    return np.concatenate([k, phi[:, None], b], axis=1)


def advect3(xyz: np.ndarray, modes: np.ndarray, dt: float, steps: int = 1) -> np.ndarray:
    """Forward Euler in the potential_modes field (host restatement of the
    device advection, for tests)."""
    p = np.array(xyz, np.float64, copy=True)
    for _ in range(steps):
        u = np.zeros_like(p)
        for kx, ky, kz, phi, bx, by, bz in modes:
            c = np.cos(kx * p[:, 0] + ky * p[:, 1] + kz * p[:, 2] + phi)
            u[:, 0] += c * bx
            u[:, 1] += c * by
            u[:, 2] += c * bz
        p += dt * u
    return p
