"""Seeded synthetic inputs for the benchmark problems.

All randomness comes from ``numpy.random.default_rng(seed)`` with the master
seed 20261202 and ``seed = 20261202 + 1000 * problem``; draws are taken in a
fixed order so that every host produces identical inputs.  Coefficient and
forcing fields are evaluated at the quadrature points in the layouts the
library expects (see include/dunefem_amd.h):

* volume points  ``[E][m^d]``, element index x fastest, point index dim 0 fastest;
* face points    per normal direction, faces on the element grid extended by
  one layer in that direction, tangential point index lower dim fastest.

Problems:

0. 2D  -Lap u + u = f, Q2 continuous, 64x64, f = 8 Fourier modes.
1. 3D  SIPG Q3 advection-diffusion, 64^3, kappa = clip(exp(g), 0.1, 10) with g
   16 Fourier modes N(0, 0.5^2), b = (1, 0.5, 0.25), beta = 10 p^2, u ~ N(0,1).
2. 256 x 2D  -div(kappa grad u) = f, Q2, side ~ U{32..256}, kappa log-normal
   (sigma 1, correlation length 0.1, random Fourier features).
3. 3D  SIPG Q2 diffusion, 48^3, kappa = 10^(4H), H binary channels (20 %).
4. 3D  SIPG Q4 diffusion, 96^3, kappa log-normal sigma 1.5 per cell.
"""
import numpy as np

from . import gauss_points

MASTER_SEED = 20261202


def seed_for(problem):
    return MASTER_SEED + 1000 * int(problem)


def fourier_modes(rng, nmodes, dim, kmax=4, sigma=1.0):
    """nmodes pairs (k in {1..kmax}^dim, amplitude ~ N(0, sigma^2)), drawn mode by mode."""
    modes = []
    for _ in range(nmodes):
        k = rng.integers(1, kmax + 1, size=dim)
        a = rng.normal(0.0, sigma)
        modes.append((k, a))
    return modes


def _sine_tables(modes, coords):
    """per mode and dimension: sin(k pi x) on the coordinate list of that dim"""
    return [np.stack([np.sin(k[d] * np.pi * coords[d]) for k, _ in modes]) for d in range(len(coords))]


def sine_sum_grid(modes, coords):
    """sum_m a_m prod_d sin(k_md pi x_d) on the tensor grid coords (list per dim);
    returns array indexed [x_{d-1}, ..., x_0] (last dim first)."""
    amps = np.array([a for _, a in modes])
    T = _sine_tables(modes, coords)
    if len(coords) == 2:
        return (T[1].T * amps) @ T[0]
    out = np.empty((len(coords[2]), len(coords[1]), len(coords[0])))
    A = T[1].T * amps  # (ny, M)
    for z in range(len(coords[2])):
        out[z] = (A * T[2][:, z]) @ T[0]
    return out


def qp_coords(n, m):
    x, _ = gauss_points(m)
    return ((np.arange(n)[:, None] + x[None, :]) / n).ravel()


def face_coords(n):
    return np.arange(n + 1) / n


def vol_layout3(G, n, m):
    nx, ny, nz = n
    G = G.reshape(nz, m, ny, m, nx, m).transpose(0, 2, 4, 1, 3, 5)
    return np.ascontiguousarray(G.reshape(nx * ny * nz, m ** 3))


def vol_layout2(G, n, m):
    nx, ny = n
    G = G.reshape(ny, m, nx, m).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(G.reshape(nx * ny, m * m))


def face_layouts3(fn, n, m):
    """fn(coords_x, coords_y, coords_z) -> grid [z, y, x]; returns 3 face arrays"""
    nx, ny, nz = n
    qx, qy, qz = qp_coords(nx, m), qp_coords(ny, m), qp_coords(nz, m)
    fx, fy, fz = face_coords(nx), face_coords(ny), face_coords(nz)
    Gx = fn(fx, qy, qz).reshape(nz, m, ny, m, nx + 1).transpose(0, 2, 4, 1, 3)
    Gy = fn(qx, fy, qz).reshape(nz, m, ny + 1, nx, m).transpose(0, 2, 3, 1, 4)
    Gz = fn(qx, qy, fz).reshape(nz + 1, ny, m, nx, m).transpose(0, 1, 3, 2, 4)
    return [np.ascontiguousarray(G.reshape(-1, m * m)) for G in (Gx, Gy, Gz)]


def rff_field(rng, nmodes, dim, ell):
    """unit-variance Gaussian field by random Fourier features: returns (W, phi)"""
    W = rng.normal(0.0, 1.0 / ell, size=(nmodes, dim))
    phi = rng.uniform(0.0, 2 * np.pi, size=nmodes)
    return W, phi


def rff_grid(W, phi, coords):
    """sqrt(2/M) sum_m cos(w_m . x + phi_m) on a tensor grid, indexed [x_{d-1},...,x_0]"""
    M = W.shape[0]
    dim = len(coords)
    E = [np.exp(1j * np.outer(W[:, d], coords[d])) for d in range(dim)]  # (M, n_d)
    c = np.exp(1j * phi)
    if dim == 2:
        G = ((E[1].T * c) @ E[0]).real
    else:
        G = np.empty((len(coords[2]), len(coords[1]), len(coords[0])))
        A = E[1].T * c
        for z in range(len(coords[2])):
            G[z] = ((A * E[2][:, z]) @ E[0]).real
    return np.sqrt(2.0 / M) * G


# ---------------------------------------------------------------------------
# problem 0 / 2: continuous Q2 in 2D
# ---------------------------------------------------------------------------
def problem0(n=64, seed=None):
    rng = np.random.default_rng(seed_for(0) if seed is None else seed)
    modes = fourier_modes(rng, 8, 2)
    m = 3
    G = sine_sum_grid(modes, [qp_coords(n, m), qp_coords(n, m)])
    return {"n": (n, n), "c": 1.0, "kappa": 1.0, "fq": vol_layout2(G, (n, n), m), "modes": modes}


def problem2_instances(count=256, smin=32, smax=256, seed=None, sigma=1.0, ell=0.1, rff_modes=128):
    rng = np.random.default_rng(seed_for(2) if seed is None else seed)
    sides = rng.integers(smin, smax + 1, size=count)
    out = []
    m = 3
    for s in sides:
        s = int(s)
        W, phi = rff_field(rng, rff_modes, 2, ell)
        modes = fourier_modes(rng, 8, 2)
        c = qp_coords(s, m)
        kq = np.exp(sigma * rff_grid(W, phi, [c, c]))
        fq = sine_sum_grid(modes, [c, c])
        out.append({"n": (s, s), "kvol": vol_layout2(kq, (s, s), m), "fq": vol_layout2(fq, (s, s), m)})
    return out


# ---------------------------------------------------------------------------
# problem 1: SIPG Q3 advection-diffusion apply
# ---------------------------------------------------------------------------
def problem1(n=64, p=3, seed=None, with_u=True):
    rng = np.random.default_rng(seed_for(1) if seed is None else seed)
    modes = fourier_modes(rng, 16, 3, sigma=0.5)
    m = p + 1
    nn = (n, n, n) if np.isscalar(n) else tuple(n)

    def kap(cx, cy, cz):
        return np.clip(np.exp(sine_sum_grid(modes, [cx, cy, cz])), 0.1, 10.0)

    kvol = vol_layout3(kap(qp_coords(nn[0], m), qp_coords(nn[1], m), qp_coords(nn[2], m)), nn, m)
    kfaces = face_layouts3(kap, nn, m)
    E = nn[0] * nn[1] * nn[2]
    u = rng.normal(0.0, 1.0, size=E * m ** 3) if with_u else None
    return {"n": nn, "p": p, "beta": 10.0 * p * p, "b": (1.0, 0.5, 0.25), "c": 0.0,
            "kvol": kvol, "kfaces": kfaces, "u": u}


# ---------------------------------------------------------------------------
# problems 3 / 4: SIPG diffusion with cellwise kappa
# ---------------------------------------------------------------------------
def channel_field(rng, n, frac=0.2):
    H = np.zeros((n, n, n), dtype=np.int8)  # [z, y, x]
    while H.mean() < frac:
        axis = int(rng.integers(0, 3))
        width = int(rng.integers(1, 4))
        i0, j0 = rng.integers(0, n, size=2)
        sl = [slice(None)] * 3
        others = [a for a in range(3) if a != axis]
        sl[2 - others[0]] = slice(int(i0), int(i0) + width)
        sl[2 - others[1]] = slice(int(j0), int(j0) + width)
        H[tuple(sl)] = 1
    return H


def problem3(n=48, p=2, seed=None, nmodes=16):
    rng = np.random.default_rng(seed_for(3) if seed is None else seed)
    H = channel_field(rng, n)
    kcell = (10.0 ** (4.0 * H)).ravel().astype(np.float64)
    modes = fourier_modes(rng, nmodes, 3)
    m = p + 1
    c = qp_coords(n, m)
    fq = vol_layout3(sine_sum_grid(modes, [c, c, c]), (n, n, n), m)
    return {"n": (n, n, n), "p": p, "beta": 10.0 * p * p, "kcell": kcell, "fq": fq, "H": H}


def problem4(n=96, p=4, seed=None, sigma=1.5, ell=0.1, rff_modes=128, nmodes=32):
    rng = np.random.default_rng(seed_for(4) if seed is None else seed)
    W, phi = rff_field(rng, rff_modes, 3, ell)
    cc = (np.arange(n) + 0.5) / n
    kcell = np.exp(sigma * rff_grid(W, phi, [cc, cc, cc])).ravel()
    modes = fourier_modes(rng, nmodes, 3)
    m = p + 1
    c = qp_coords(n, m)
    fq = vol_layout3(sine_sum_grid(modes, [c, c, c]), (n, n, n), m)
    return {"n": (n, n, n), "p": p, "beta": 10.0 * p * p, "kcell": kcell, "fq": fq}
