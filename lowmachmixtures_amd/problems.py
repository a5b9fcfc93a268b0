"""Synthetic Stokes problems and the host-side Gaussian noise generator.

Inputs for the solver are generated once on the host from integer seeds and
then shared, bit for bit, between the device path and any checker.

Noise generator: MT19937 seeded with the 2002 ``init_genrand`` rule and
P. J. Acklam's inverse-normal rational approximation, the two published
methods behind the reference's ``genrandn`` (src_random/RNGs.c:62-260).
Uniforms are ``u32 * 2**-32``; a zero uniform is redrawn.  numpy's MT19937
with legacy 32-bit seeding produces the same raw stream.

Field layout: every array is a deduplicated periodic field of shape
``(nz, ny, nx)`` (``(ny, nx)`` in 2D), x fastest.  Face component d stores
the lower face of each cell in direction d.  In 2D, ``beta_ed`` holds one
node array (node (i,j) = lower-left corner of cell (i,j)); in 3D it holds
the xy, xz and yz edge arrays.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Acklam's coefficients (published with the method)
_A = (-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
      1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00)
_B = (-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
      6.680131188771972e+01, -1.328068155288572e+01)
_C = (-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
      -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00)
_D = (7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
      3.754408661907416e+00)


def mt_raw(seed: int, count: int) -> np.ndarray:
    """First ``count`` raw 32-bit MT19937 outputs after init_genrand(seed)."""
    bg = np.random.MT19937()
    bg._legacy_seeding(int(seed) & 0xFFFFFFFF)
    return bg.random_raw(int(count)).astype(np.uint32)


def inv_norm_dist(p: np.ndarray) -> np.ndarray:
    """Acklam's approximation of the inverse standard normal CDF (|rel err| < 1.15e-9)."""
    p = np.asarray(p, dtype=np.float64)
    x = np.empty_like(p)
    plow, phigh = 0.02425, 1.0 - 0.02425
    lo = p < plow
    hi = p > phigh
    mid = ~(lo | hi)
    q = p[mid] - 0.5
    r = q * q
    a, b = _A, _B
    x[mid] = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q / \
             (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0)
    c, d = _C, _D
    for mask, sgn, pv in ((lo, 1.0, p[lo]), (hi, -1.0, 1.0 - p[hi])):
        if not mask.any():
            continue
        q = np.sqrt(-2.0 * np.log(pv))
        num = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5])
        den = ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0)
        x[mask] = num / den if sgn > 0 else -num / den
    return x


def uniforms(seed: int, count: int) -> np.ndarray:
    """U[0,1) stream of the generator (zeros kept)."""
    return mt_raw(seed, count).astype(np.float64) * (1.0 / 4294967296.0)


def normals(seed: int, count: int) -> np.ndarray:
    """``count`` N(0,1) samples: genrandn's stream, zero uniforms redrawn."""
    count = int(count)
    raw = mt_raw(seed, count + 64)
    u = raw.astype(np.float64) * (1.0 / 4294967296.0)
    u = u[u != 0.0]
    extra = 64
    while u.size < count:  # astronomically unlikely
        extra *= 4
        raw = mt_raw(seed, count + extra)
        u = raw.astype(np.float64) * (1.0 / 4294967296.0)
        u = u[u != 0.0]
    return inv_norm_dist(u[:count])


# ---------------------------------------------------------------------------
# periodic averaging helpers (convert_stag.f90 interior formulas)
# ---------------------------------------------------------------------------

def _roll(a, shift, axis):
    return np.roll(a, shift, axis=axis)


def box_filter(a: np.ndarray, passes: int = 1) -> np.ndarray:
    """Periodic 3^d box mean, applied ``passes`` times (as d separable 3-point means)."""
    for _ in range(passes):
        for ax in range(a.ndim):
            a = (_roll(a, 1, ax) + a + _roll(a, -1, ax)) / 3.0
    return a


def avg_cc_to_face(cc: np.ndarray):
    nd = cc.ndim
    # axis order in the array is (z, y, x); face direction d maps to axis nd-1-d
    return [0.5 * (cc + _roll(cc, 1, nd - 1 - d)) for d in range(nd)]


def avg_cc_to_node_2d(cc: np.ndarray) -> np.ndarray:
    im = _roll(cc, 1, 1)
    jm = _roll(cc, 1, 0)
    ijm = _roll(im, 1, 0)
    return 0.25 * (cc + im + jm + ijm)


def avg_cc_to_edge_3d(cc: np.ndarray):
    im, jm, km = _roll(cc, 1, 2), _roll(cc, 1, 1), _roll(cc, 1, 0)
    xy = 0.25 * (cc + im + jm + _roll(im, 1, 1))
    xz = 0.25 * (cc + im + km + _roll(im, 1, 0))
    yz = 0.25 * (cc + jm + km + _roll(jm, 1, 0))
    return [xy, xz, yz]


@dataclass
class StokesProblem:
    """One periodic variable-viscosity staggered Stokes problem."""
    dim: int
    n: tuple
    dx: float
    visc_type: int
    theta_alpha: float
    alpha: list            # dim face arrays
    beta: np.ndarray       # cell
    beta_ed: list          # 1 node array (2D) / 3 edge arrays (3D)
    gamma: np.ndarray      # cell
    W_cc: np.ndarray       # (dim, *shape) white noise on cells
    W_ed: list             # list of (2, *shape) white noise on nodes/edges
    b_p: np.ndarray
    solver: dict = field(default_factory=dict)
    name: str = ""

    @property
    def shape(self):
        return tuple(reversed(self.n[: self.dim]))

    @property
    def ncells(self):
        return int(np.prod(self.n[: self.dim]))


def _noise(seed, dim, shape):
    n = int(np.prod(shape))
    ne = 1 if dim == 2 else 3
    z = normals(seed, n * (dim + 2 * ne))
    W_cc = z[: dim * n].reshape((dim,) + shape).copy()
    W_ed = []
    off = dim * n
    for _ in range(ne):
        W_ed.append(z[off: off + 2 * n].reshape((2,) + shape).copy())
        off += 2 * n
    return W_cc, W_ed


def config1(n: int = 128, seed: int = 1) -> StokesProblem:
    """Plain FP64 2D staggered Stokes solve on one periodic n x n box, dx = 1.

    theta*rho = 1; eta = exp(Z ln10 / 2), Z ~ N(0,1), smoothed by one
    periodic 3x3 box filter; eta on nodes by 4-point averaging; gamma = 0;
    RHS = divergence of a white-noise staggered stress field (variance 1);
    GMRES tol 1e-9 with 2 V-cycles of each multigrid per preconditioner call.
    """
    shape = (n, n)
    z = normals(seed, n * n).reshape(shape)
    eta = box_filter(np.exp(z * math.log(10.0) / 2.0), 1)
    W_cc, W_ed = _noise(seed + 1, 2, shape)
    return StokesProblem(
        dim=2, n=(n, n, 1), dx=1.0, visc_type=-1, theta_alpha=1.0,
        alpha=[np.ones(shape), np.ones(shape)], beta=eta, beta_ed=[avg_cc_to_node_2d(eta)],
        gamma=np.zeros(shape), W_cc=W_cc, W_ed=W_ed, b_p=np.zeros(shape),
        solver=dict(visc_type=-1, stag_mg_max_vcycles=2, mg_max_vcycles=2, gmres_rel_tol=1e-9),
        name=f"2d_periodic_{n}x{n}_visc-1",
    )


def config3(n: int = 256, seed: int = 1, contrast: float = 100.0) -> StokesProblem:
    """3D periodic n^3 variable-viscosity Stokes solve, dx = 1.

    eta = exp(sigma Z), sigma = ln(contrast)/4, two passes of the periodic
    27-point mean; eta on edges by 4-point averaging; rho ~ U[1,4] with
    alpha on faces by 2-point averaging; theta = 1; RHS = divergence of
    white-noise stress; reference-default solver settings, tol 1e-9.
    """
    shape = (n, n, n)
    N = n ** 3
    z = normals(seed, N).reshape(shape)
    eta = box_filter(np.exp(z * (math.log(contrast) / 4.0)), 2)
    rho = 1.0 + 3.0 * uniforms(seed + 1, N).reshape(shape)
    W_cc, W_ed = _noise(seed + 2, 3, shape)
    return StokesProblem(
        dim=3, n=(n, n, n), dx=1.0, visc_type=-1, theta_alpha=1.0,
        alpha=avg_cc_to_face(rho), beta=eta, beta_ed=avg_cc_to_edge_3d(eta),
        gamma=np.zeros(shape), W_cc=W_cc, W_ed=W_ed, b_p=np.zeros(shape),
        solver=dict(visc_type=-1, gmres_rel_tol=1e-9),
        name=f"3d_periodic_{n}^3_visc-1_contrast{contrast:g}",
    )


def config4_member(n: int, seed: int, contrast: float = 10.0) -> StokesProblem:
    """One member of the batched 2D ensemble: a periodic n x n box, dx = 1.

    eta = exp(sigma Z), sigma = ln(contrast)/4 (contrast 10), one periodic
    3x3 box filter, eta on nodes by 4-point averaging; theta*rho = 1;
    gamma = 0; RHS = divergence of the member's own white-noise stress;
    reference-default solver settings (one V-cycle of each multigrid),
    tol 1e-9, FP32 V-cycles inside FP64 flexible GMRES.
    """
    shape = (n, n)
    z = normals(seed, n * n).reshape(shape)
    eta = box_filter(np.exp(z * (math.log(contrast) / 4.0)), 1)
    W_cc, W_ed = _noise(seed + 1, 2, shape)
    return StokesProblem(
        dim=2, n=(n, n, 1), dx=1.0, visc_type=-1, theta_alpha=1.0,
        alpha=[np.ones(shape), np.ones(shape)], beta=eta, beta_ed=[avg_cc_to_node_2d(eta)],
        gamma=np.zeros(shape), W_cc=W_cc, W_ed=W_ed, b_p=np.zeros(shape),
        solver=dict(visc_type=-1, gmres_rel_tol=1e-9, precision=1),
        name=f"2d_periodic_{n}x{n}_contrast{contrast:g}_seed{seed}",
    )


def config4(per_size: int = 32, sizes=(64, 128, 256), seed: int = 1):
    """The ragged ensemble: per_size members of each size, each with its own seeds."""
    return [config4_member(n, seed + 1000 * k + 2 * m) for k, n in enumerate(sizes) for m in range(per_size)]


def rhs_from_noise_host(p: StokesProblem):
    """b_u = div(W) on faces with eta = T = 1, variance 1 (stoch_m_force_*), numpy.

    Used only to build small test problems on the host; the device path
    computes the same divergence with its own kernel.
    """
    dim = p.dim
    nd = dim
    if dim == 2:
        nd1, nd2 = p.W_ed[0][0], p.W_ed[0][1]
        c1, c2 = p.W_cc[0], p.W_cc[1]
        bx = (c1 - _roll(c1, 1, 1)) + (_roll(nd1, -1, 0) - nd1)
        by = (_roll(nd2, -1, 1) - nd2) + (c2 - _roll(c2, 1, 0))
        return [bx / p.dx, by / p.dx]
    xy1, xy2 = p.W_ed[0]
    xz1, xz2 = p.W_ed[1]
    yz1, yz2 = p.W_ed[2]
    c1, c2, c3 = p.W_cc
    bx = (c1 - _roll(c1, 1, 2)) + (_roll(xy1, -1, 1) - xy1) + (_roll(xz1, -1, 0) - xz1)
    by = (_roll(xy2, -1, 2) - xy2) + (c2 - _roll(c2, 1, 1)) + (_roll(yz1, -1, 0) - yz1)
    bz = (_roll(xz2, -1, 2) - xz2) + (_roll(yz2, -1, 1) - yz2) + (c3 - _roll(c3, 1, 0))
    del nd
    return [bx / p.dx, by / p.dx, bz / p.dx]


def small_problem(dim: int, n, seed: int = 3, visc_type: int = -1, contrast: float = 10.0,
                  theta_alpha: float = 1.0) -> StokesProblem:
    """Small random problem for parity tests (any n per direction)."""
    if isinstance(n, int):
        n = (n,) * dim
    n = tuple(n) + (1,) * (3 - len(n))
    shape = tuple(reversed(n[:dim]))
    N = int(np.prod(shape))
    z = normals(seed, N).reshape(shape)
    eta = box_filter(np.exp(z * (math.log(contrast) / 4.0)), 1)
    if abs(visc_type) in (1, 2, 3) and visc_type > 0:
        eta = np.full(shape, float(eta.mean()))
    rho = 1.0 + 3.0 * uniforms(seed + 1, N).reshape(shape)
    gamma = 0.5 * eta if abs(visc_type) == 3 else np.zeros(shape)
    W_cc, W_ed = _noise(seed + 2, dim, shape)
    bed = [avg_cc_to_node_2d(eta)] if dim == 2 else avg_cc_to_edge_3d(eta)
    return StokesProblem(
        dim=dim, n=n, dx=1.0 if dim == 2 else 0.5, visc_type=visc_type, theta_alpha=theta_alpha,
        alpha=avg_cc_to_face(rho), beta=eta, beta_ed=bed, gamma=gamma, W_cc=W_cc, W_ed=W_ed,
        b_p=np.zeros(shape), solver=dict(visc_type=visc_type), name=f"small_{dim}d_{n}",
    )
