"""Initial states of the benchmark/parity workloads (synthetic, seeded).

Nodal initialisation at the Gauss-Legendre nodes of each cell, as the
reference's solutionAdjustment does (ExaHyPE/kernels/aderdg/generic/c/3d/
solutionAdjustment.cpph:18-63: x = centre + dx * (xi - 1/2)).  Layout
luh[cell][z][y][x][var], cells x fastest, unit cube / square, n cells per side.

* density_wave_2d  -- 2D Euler, rho = 1 + 0.2 sin(2 pi (x + y)), u = v = 1, p = 1
* density_wave_3d  -- 3D Euler, rho = 1 + 0.2 sin(2 pi (x + y + z)) + sum of
                      M seeded Fourier modes of amplitude 0.01/M (|k_i| <= 4) in
                      rho and in the x-velocity u; v = w = 1, p = 1
* amr_* / ensemble_* -- batched uneven work: a static two-level 3D mesh (27^3
                      coarse cells, those with centre within 0.25 of the cube
                      centre split 3x) and 64 independent 2D meshes with seeded
                      sizes between 9 and 81 cells per side
* acoustic_*       -- 3D linear acoustics: per-cell wave speed
                      c = 1 + sum of 8 seeded Gaussian bumps (+-0.2, width 0.1),
                      rho = 1, a Ricker point source (f0 = 5) and 64 seeded
                      receivers around it
"""
import numpy as np

GAMMA = 1.4


def _nodes(order):
    x, _ = np.polynomial.legendre.leggauss(order + 1)
    return (x + 1.0) / 2.0


def fourier_modes(seed=0, modes=8, kmax=4):
    """(k[M,3], phase[M]) for rho and for u, drawn from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(2):
        ks = []
        while len(ks) < modes:
            k = rng.integers(-kmax, kmax + 1, size=3)
            if np.any(k != 0):
                ks.append(k)
        ph = rng.uniform(0.0, 2.0 * np.pi, size=modes)
        out.append((np.array(ks, dtype=np.float64), ph))
    return out


def _coords(n, order, dim, xp, device):
    h = 1.0 / n
    if xp is np:
        xi = _nodes(order)
        ar = np.arange(n, dtype=np.float64)
    else:
        xi = xp.as_tensor(_nodes(order), dtype=xp.float64, device=device)
        ar = xp.arange(n, dtype=xp.float64, device=device)
    centres = (ar + 0.5) * h
    # 1D node coordinates per cell index: [n][B]
    x1 = centres[:, None] + h * (xi[None, :] - 0.5)
    return x1, h


def density_wave_2d(n, order, xp=np, device=None):
    x1, h = _coords(n, order, 2, xp, device)
    B = order + 1
    # cell (iy, ix), node (jy, jx)
    X = x1[None, :, None, :]                  # [1, ix, 1, jx]
    Y = x1[:, None, :, None]                  # [iy, 1, jy, 1]
    rho = 1.0 + 0.2 * xp.sin(2 * np.pi * (X + Y))
    rho = rho + 0 * (X + Y)
    u = v = 1.0
    p = 1.0
    Q = xp.zeros((n, n, B, B, 5), dtype=xp.float64, **({} if xp is np else {"device": device}))
    Q[..., 0] = rho
    Q[..., 1] = rho * u
    Q[..., 2] = rho * v
    Q[..., 4] = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    # [iy, ix, jy, jx, v] -> [cell=(iy, ix)][jy][jx][v]
    return Q.reshape(-1)


def density_wave_3d(n, order, seed=0, modes=8, xp=np, device=None):
    x1, h = _coords(n, order, 3, xp, device)
    B = order + 1
    (kr, pr), (ku, pu) = fourier_modes(seed, modes)
    kw = {} if xp is np else {"device": device}
    # broadcast shapes: [iz, iy, ix, jz, jy, jx]
    X = x1.reshape(1, 1, n, 1, 1, B)
    Y = x1.reshape(1, n, 1, 1, B, 1)
    Z = x1.reshape(n, 1, 1, B, 1, 1)
    rho = 1.0 + 0.2 * xp.sin(2 * np.pi * (X + Y + Z))
    u = 1.0 + 0.0 * (X + Y + Z)
    amp = 0.01 / modes if modes else 0.0
    for m in range(modes):
        rho = rho + amp * xp.cos(2 * np.pi * (kr[m, 0] * X + kr[m, 1] * Y + kr[m, 2] * Z) + pr[m])
        u = u + amp * xp.cos(2 * np.pi * (ku[m, 0] * X + ku[m, 1] * Y + ku[m, 2] * Z) + pu[m])
    v = w = 1.0
    p = 1.0
    Q = xp.zeros((n, n, n, B, B, B, 5), dtype=xp.float64, **kw)
    Q[..., 0] = rho
    Q[..., 1] = rho * u
    Q[..., 2] = rho * v
    Q[..., 3] = rho * w
    Q[..., 4] = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return Q.reshape(-1)


def density_wave_params(dim, seed=0, modes=8):
    """Parameter vector of the device initial condition EXA_IC_DENSITY_WAVE that
    reproduces density_wave_2d (dim 2) / density_wave_3d (dim 3)."""
    if dim == 2:
        return np.array([0.2, 1.0, 1.0, 0.0, 1.0, 0, 0.0])
    (kr, pr), (ku, pu) = fourier_modes(seed, modes)
    prm = [0.2, 1.0, 1.0, 1.0, 1.0, modes, 0.01 / modes if modes else 0.0]
    for k, ph in ((kr, pr), (ku, pu)):
        for m in range(modes):
            prm += [k[m, 0], k[m, 1], k[m, 2], ph[m]]
    return np.array(prm, dtype=np.float64)


def dof_per_cell(dim, order, nvar=5):
    return (order + 1) ** dim * nvar


# ---- linear acoustics ---------------------------------------------------------
ACOUSTIC_SOURCE = (0.5, 0.5, 0.5)
ACOUSTIC_F0 = 5.0
ACOUSTIC_T0 = 0.06
ACOUSTIC_AMP = 1.0


def acoustic_bumps(seed=0, nbumps=8, amp=0.2, width=0.1):
    """(centres[nb,3], signed amplitudes[nb]) of the wave-speed bumps."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, size=(nbumps, 3))
    signs = rng.choice([-1.0, 1.0], size=nbumps)
    return centres, amp * signs, width


def acoustic_wave_speed(n, seed=0, scale=1.0, nbumps=8):
    """c per cell (cells x fastest) = 1 + scale * sum_b a_b exp(-|x - x_b|^2 / (2 w^2)),
    evaluated at the cell centres of the unit cube."""
    centres, amps, width = acoustic_bumps(seed, nbumps)
    ar = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(ar, ar, ar, indexing="ij")
    pts = np.stack([X, Y, Z], -1).reshape(-1, 3)
    c = np.ones(len(pts))
    for xb, a in zip(centres, amps):
        c += scale * a * np.exp(-np.sum((pts - xb) ** 2, axis=1) / (2 * width ** 2))
    return c


def acoustic_receivers(seed=0, nrec=64, centre=ACOUSTIC_SOURCE, radius=0.12):
    """nrec receiver positions: a 4x4x4 lattice of spacing radius*2/3 around
    the source, each jittered by up to 1/10 of the spacing (seeded)."""
    rng = np.random.default_rng(seed + 1000)
    m = int(round(nrec ** (1 / 3)))
    assert m ** 3 == nrec
    g = np.linspace(-radius, radius, m)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X, Y, Z], -1).reshape(-1, 3) + np.asarray(centre)
    pts += rng.uniform(-0.1, 0.1, size=pts.shape) * (2 * radius / (m - 1))
    return pts


def acoustic_state(n, order, c, rho=None, u=None):
    """Reference-layout state luh[cell][B^3][6] = (p,u,v,w,rho,c) with the given
    per-cell material and variables (zero if u is None)."""
    B3 = (order + 1) ** 3
    nc = n ** 3
    luh = np.zeros((nc, B3, 6))
    if u is not None:
        luh[..., :4] = u
    luh[..., 4] = (np.ones(nc) if rho is None else np.asarray(rho))[:, None]
    luh[..., 5] = np.asarray(c)[:, None]
    return luh


# ---- batched uneven work ---------------------------------------------------------
def amr_ball_refinement(n=27, radius=0.25, centre=(0.5, 0.5, 0.5)):
    """refined[c] = 1 for coarse cells (x fastest) whose centre lies within radius."""
    ar = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(ar, ar, ar, indexing="ij")
    d2 = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
    return (d2.ravel() <= radius ** 2).astype(np.uint8)


def amr_leaf_points(order, n, level, coords):
    """Node coordinates [nleaves][B^3][3] of the leaves of a two-level mesh."""
    x = _nodes(order)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    loc = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    h = np.where(np.asarray(level) == 0, 1.0 / n, 1.0 / (3 * n))
    return (np.asarray(coords)[:, None, :] + loc[None]) * h[:, None, None]


def node_points(n, order, dim):
    """Node coordinates [n^D cells][B^D nodes][dim] of the uniform mesh on the
    unit box, cells and nodes x fastest (x = centre + h (xi - 1/2))."""
    x1, _ = _coords(n, order, dim, np, None)  # [n][B]
    B = order + 1
    if dim == 2:
        sh = (n, n, B, B)
        X = np.broadcast_to(x1[None, :, None, :], sh)
        Y = np.broadcast_to(x1[:, None, :, None], sh)
        return np.stack([X, Y], -1).reshape(n * n, B * B, 2)
    sh = (n, n, n, B, B, B)
    X = np.broadcast_to(x1[None, None, :, None, None, :], sh)
    Y = np.broadcast_to(x1[None, :, None, None, :, None], sh)
    Z = np.broadcast_to(x1[:, None, None, :, None, None], sh)
    return np.stack([X, Y, Z], -1).reshape(n ** 3, B ** 3, 3)


def euler_state(rho, u, v, w, p):
    return np.stack([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)], -1)


def amr_density_wave(order, n, level, coords, amp=0.2):
    """3D density wave rho = 1 + amp sin(2 pi (x+y+z)), u = v = w = 1, p = 1 on the leaves."""
    pts = amr_leaf_points(order, n, level, coords)
    rho = 1.0 + amp * np.sin(2 * np.pi * (pts[..., 0] + pts[..., 1] + pts[..., 2]))
    one = np.ones_like(rho)
    return np.ascontiguousarray(euler_state(rho, one, one, one, one))


def ensemble_sizes(seed=0, count=64, lo=9, hi=81):
    """Seeded cells per side of the independent 2D instances."""
    return np.random.default_rng(seed).integers(lo, hi + 1, size=count).astype(np.int32)
