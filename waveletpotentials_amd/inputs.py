"""Seeded synthetic inputs for the recovery workloads (host side, float64 numpy).

Velocities are the discrete curl of a seeded potential whose boundary-tangential
components are zeroed (the zero-normal-flux construction of
WaveletPotentials/StreamInitial.cpp:5-85), so they are discretely
divergence-free and satisfy velocity bc 'n'.  The discrete curl follows
StreamInitial::GradientStream (WaveletPotentials/StreamInitial.cpp:87-175).

Randomness comes from numpy's PCG64 bit generator with an explicit seed, so the
inputs are the same on every machine.  Arrays use the (nz, nx, ny) layout.
"""
import numpy as np

__all__ = ["seed_for", "spectral_potential3d", "curl3d", "curl2d", "div3d", "velocity3d",
           "gaussian_vortex_stream2d", "velocity2d", "spheres", "BATCH_SIZES", "batch_problems"]


def seed_for(config_id, problem_id=0):
    return 20260001 + 1000 * config_id + problem_id


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _modes(kmax, rng):
    r = np.arange(-kmax, kmax + 1)
    kx, ky, kz = np.meshgrid(r, r, r, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    mask = (k2 >= 1) & (k2 <= kmax * kmax)
    kmag = np.sqrt(np.where(mask, k2, 1.0))
    amp = np.where(mask, kmag ** (-11.0 / 6.0), 0.0)
    shape = kx.shape
    phase = rng.uniform(0.0, 2.0 * np.pi, size=shape)
    pol = rng.standard_normal(size=shape + (3,))
    pol /= np.linalg.norm(pol, axis=-1, keepdims=True)
    return r, amp, phase, pol


def _eval_component(coef, r, xs, ys, zs):
    """Re sum_k coef[kx,ky,kz] e^{2 pi i (kx x + ky y + kz z)} on the tensor grid
    (zs, xs, ys) -> array (len(zs), len(xs), len(ys))."""
    ex = np.exp(2j * np.pi * np.outer(xs, r))  # (nx, K)
    ey = np.exp(2j * np.pi * np.outer(ys, r))
    ez = np.exp(2j * np.pi * np.outer(zs, r))
    # contract kz, then ky, then kx
    t = np.tensordot(coef, ez, axes=([2], [1]))       # (Kx, Ky, nz)
    t = np.tensordot(t, ey, axes=([1], [1]))          # (Kx, nz, ny)
    t = np.tensordot(ex, t, axes=([1], [0]))          # (nx, nz, ny)
    return np.ascontiguousarray(np.transpose(t.real, (1, 0, 2)))


def spectral_potential3d(n, kmax, seed, extent=(1.0, 1.0, 1.0)):
    """Edge potential (qx, qy, qz) from Fourier modes 1 <= |k| <= kmax with random
    phases, random unit polarisation and amplitude |k|^(-11/6); boundary-
    tangential components are zeroed."""
    nx, ny, nz = n
    hx, hy, hz = extent[0] / nx, extent[1] / ny, extent[2] / nz
    rng = _rng(seed)
    r, amp, phase, pol = _modes(kmax, rng)
    base = amp * np.exp(1j * phase)
    xc = (np.arange(nx) + 0.5) * hx / extent[0]
    yc = (np.arange(ny) + 0.5) * hy / extent[1]
    zc = (np.arange(nz) + 0.5) * hz / extent[2]
    xn = np.arange(nx + 1) * hx / extent[0]
    yn = np.arange(ny + 1) * hy / extent[1]
    zn = np.arange(nz + 1) * hz / extent[2]
    qx = _eval_component(base * pol[..., 0], r, xc, yn, zn)  # (nz+1, nx, ny+1)
    qy = _eval_component(base * pol[..., 1], r, xn, yc, zn)  # (nz+1, nx+1, ny)
    qz = _eval_component(base * pol[..., 2], r, xn, yn, zc)  # (nz, nx+1, ny+1)
    qx[0] = qx[-1] = 0.0
    qx[:, :, 0] = qx[:, :, -1] = 0.0
    qy[0] = qy[-1] = 0.0
    qy[:, 0, :] = qy[:, -1, :] = 0.0
    qz[:, 0, :] = qz[:, -1, :] = 0.0
    qz[:, :, 0] = qz[:, :, -1] = 0.0
    return qx, qy, qz


def curl3d(qx, qy, qz, dx):
    """Face velocities (xv, yv, zv) = curl of the edge potential."""
    hx, hy, hz = dx
    xv = (qz[:, :, 1:] - qz[:, :, :-1]) / hy - (qy[1:] - qy[:-1]) / hz
    yv = (qx[1:] - qx[:-1]) / hz - (qz[:, 1:, :] - qz[:, :-1, :]) / hx
    zv = (qy[:, 1:, :] - qy[:, :-1, :]) / hx - (qx[:, :, 1:] - qx[:, :, :-1]) / hy
    return (np.ascontiguousarray(xv), np.ascontiguousarray(yv), np.ascontiguousarray(zv))


def div3d(xv, yv, zv, dx):
    hx, hy, hz = dx
    return ((xv[:, 1:, :] - xv[:, :-1, :]) / hx + (yv[:, :, 1:] - yv[:, :, :-1]) / hy
            + (zv[1:] - zv[:-1]) / hz)


def velocity3d(n, kmax, seed):
    """Divergence-free, zero-flux face velocity on the unit cube (n cells per axis)."""
    dx = tuple(1.0 / v for v in n)
    q = spectral_potential3d(n, kmax, seed)
    return curl3d(*q, dx), q, dx


def gaussian_vortex_stream2d(n, seed, count=16):
    """Node stream function: sum of `count` Gaussian vortices, centres U[0.1,0.9]^2,
    radii U[0.02,0.1], strengths N(0,1); zero on the boundary ring.
    Returns (nx+1, ny+1)."""
    nx, ny = n
    rng = _rng(seed)
    c = rng.uniform(0.1, 0.9, size=(count, 2))
    rad = rng.uniform(0.02, 0.1, size=count)
    s = rng.standard_normal(count)
    x = np.arange(nx + 1)[:, None] / nx
    y = np.arange(ny + 1)[None, :] / ny
    psi = np.zeros((nx + 1, ny + 1))
    for v in range(count):
        psi += s[v] * np.exp(-((x - c[v, 0]) ** 2 + (y - c[v, 1]) ** 2) / (2 * rad[v] ** 2))
    psi[0, :] = psi[-1, :] = 0.0
    psi[:, 0] = psi[:, -1] = 0.0
    return psi


def curl2d(qz, dx):
    hx, hy = dx
    xv = (qz[:, 1:] - qz[:, :-1]) / hy
    yv = -(qz[1:, :] - qz[:-1, :]) / hx
    return np.ascontiguousarray(xv), np.ascontiguousarray(yv)


def velocity2d(n, seed):
    dx = tuple(1.0 / v for v in n)
    psi = gaussian_vortex_stream2d(n, seed)
    return curl2d(psi, dx), psi, dx


def spheres(count, seed):
    """`count` spheres of the solid-obstacle workloads: centres U[0.2, 0.8]^3,
    radii U[0.05, 0.15] (unit cube).  Returns (centres (count, 3), radii)."""
    rng = _rng(seed)
    return rng.uniform(0.2, 0.8, size=(count, 3)), rng.uniform(0.05, 0.15, size=count)


# extents (nx, ny, nz) of the batched workload's problems
BATCH_SIZES = ((32, 32, 32), (64, 64, 64), (128, 128, 128), (128, 64, 64))


def batch_problems(count=64, seed=None):
    """The batched workload: `count` problems with extents drawn uniformly from
    BATCH_SIZES, each with a spectral velocity (modes |k| <= min(extent) / 8 per
    box, |k|^(-11/6) amplitudes) and 0-4 spheres.  Cells are cubes of side
    1 / max(extent), so the 128x64x64 grids cover a 1 x 0.5 x 0.5 box (the
    multigrid's rediscretized coarse operator with piecewise-constant
    prolongation diverges at a 2:1 cell aspect); sphere centres are drawn in
    box-relative coordinates.  Returns a list of dicts: dims, dx, extent, kmax,
    seed, centers, radii."""
    rng = _rng(seed_for(4) if seed is None else seed)
    out = []
    for i in range(count):
        d = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        h = 1.0 / max(d)
        ext = tuple(v * h for v in d)
        c, r = spheres(int(rng.integers(0, 5)), seed_for(4, 1000 + i))
        out.append({"dims": d, "dx": (h, h, h), "extent": ext, "kmax": max(2, min(d) // 8),
                    "seed": seed_for(4, i), "centers": c * np.array(ext), "radii": r})
    return out
