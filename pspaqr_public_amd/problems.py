"""Synthetic test matrices and right-hand sides for the spaQR path.

Every generator returns a square scipy CSC matrix in natural tensor-grid ordering
(id = sum_i x_i * n^(d-1-i), the ordering of ``linspace_nd`` in the reference,
src/util.cpp:901-930), so the geometric partition ``-n n -d d`` applies to it.

Random streams reproduce ``std::mt19937`` + ``std::uniform_real_distribution<double>``
bit for bit (the reference's ``random()``, src/util.cpp:992-1001): numpy's legacy
``RandomState(seed)`` seeds MT19937 with the same ``init_genrand`` and its raw 32-bit
outputs are combined the way libstdc++'s ``generate_canonical<double, 53>`` does.
Normal variates use Box-Muller on those canonical draws (portable across standard
libraries, unlike ``std::normal_distribution``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import os

import numpy as np
import scipy.sparse as sp

_TWO32 = 4294967296.0
_TWO64 = 18446744073709551616.0


def mt_canonical(seed: int, count: int) -> np.ndarray:
    """``count`` draws of generate_canonical<double,53>(std::mt19937(seed))."""
    rs = np.random.RandomState(seed)
    raw = rs._bit_generator.random_raw(2 * count).astype(np.float64)
    s = raw[0::2] + raw[1::2] * _TWO32
    u = s / _TWO64
    return np.minimum(u, np.nextafter(1.0, 0.0))


def mt_uniform(seed: int, count: int, a: float = -1.0, b: float = 1.0) -> np.ndarray:
    """std::uniform_real_distribution<double>(a, b) over std::mt19937(seed)."""
    return mt_canonical(seed, count) * (b - a) + a


def mt_normal(seed: int, count: int, sigma: float = 1.0) -> np.ndarray:
    """N(0, sigma^2) by Box-Muller on canonical mt19937 draws (pairs -> two variates)."""
    npairs = (count + 1) // 2
    u = mt_canonical(seed, 2 * npairs)
    u1, u2 = u[0::2], u[1::2]
    rad = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.empty(2 * npairs)
    z[0::2] = rad * np.cos(2.0 * math.pi * u2)
    z[1::2] = rad * np.sin(2.0 * math.pi * u2)
    return sigma * z[:count]


def _grid_index(n: int, d: int):
    idx = np.arange(n ** d).reshape((n,) * d)
    return idx


def _assemble(n: int, d: int, face_coef, diag_extra=None, upwind=None) -> sp.csc_matrix:
    """7/5-point operator: sum over axes of -kappa_face (u_nbr - u_i); Dirichlet boundary.

    face_coef(axis) -> array of shape grid with the coefficient of the face between cell i
    and cell i + e_axis (last index along the axis is the boundary face to the right);
    the left boundary face of cell 0 uses face_coef_left(axis).
    """
    N = n ** d
    idx = _grid_index(n, d)
    rows, cols, vals = [], [], []
    diag = np.zeros((n,) * d)
    for ax in range(d):
        kf, kl = face_coef(ax)  # kf: right faces (n along ax incl. boundary), kl: left boundary face
        # interior faces between i and i+e_ax: slices [0:n-1] along ax
        sl_a = [slice(None)] * d
        sl_b = [slice(None)] * d
        sl_a[ax] = slice(0, n - 1)
        sl_b[ax] = slice(1, n)
        k_int = kf[tuple(sl_a)]
        ia = idx[tuple(sl_a)].ravel()
        ib = idx[tuple(sl_b)].ravel()
        kv = k_int.ravel()
        rows += [ia, ib]
        cols += [ib, ia]
        vals += [-kv, -kv]
        diag[tuple(sl_a)] += k_int
        diag[tuple(sl_b)] += k_int
        # boundary faces
        sl_last = [slice(None)] * d
        sl_last[ax] = slice(n - 1, n)
        diag[tuple(sl_last)] += kf[tuple(sl_last)]
        sl_first = [slice(None)] * d
        sl_first[ax] = slice(0, 1)
        diag[tuple(sl_first)] += kl
    if upwind is not None:
        # first-order upwind convection b . grad u with b_ax >= 0: b_ax (u_i - u_{i-e_ax})
        for ax, bax in enumerate(upwind):
            if bax == 0.0:
                continue
            diag += bax
            sl_a = [slice(None)] * d
            sl_b = [slice(None)] * d
            sl_a[ax] = slice(1, n)
            sl_b[ax] = slice(0, n - 1)
            ia = idx[tuple(sl_a)].ravel()
            ib = idx[tuple(sl_b)].ravel()
            rows.append(ia)
            cols.append(ib)
            vals.append(np.full(ia.size, -bax))
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A = A.tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


def poisson(n: int, d: int) -> sp.csc_matrix:
    """5-point (d=2) / 7-point (d=3) Poisson, Dirichlet: A_ii = 2d, A_ij = -1."""
    ones = np.ones((n,) * d)
    return _assemble(n, d, lambda ax: (ones, ones[tuple([slice(0, 1) if a == ax else slice(None) for a in range(d)])]))


def _harmonic_faces(kcell: np.ndarray, d: int, axis_scale=None):
    n = kcell.shape[0]

    def face(ax):
        s = 1.0 if axis_scale is None else axis_scale[ax]
        kf = np.empty_like(kcell)
        sl_a = [slice(None)] * d
        sl_b = [slice(None)] * d
        sl_a[ax] = slice(0, n - 1)
        sl_b[ax] = slice(1, n)
        a, b = kcell[tuple(sl_a)], kcell[tuple(sl_b)]
        kf[tuple(sl_a)] = 2.0 * a * b / (a + b)
        sl_last = [slice(None)] * d
        sl_last[ax] = slice(n - 1, n)
        kf[tuple(sl_last)] = kcell[tuple(sl_last)]  # boundary face uses the cell value
        sl_first = [slice(None)] * d
        sl_first[ax] = slice(0, 1)
        return s * kf, s * kcell[tuple(sl_first)]

    return face


def diffusion_varcoef(n: int, d: int, seed: int, sigma: float = 1.0, axis_scale=None) -> sp.csc_matrix:
    """Cell coefficient exp(N(0, sigma^2)); face value = harmonic mean of its two cells;
    boundary faces take the cell value; optional per-axis conductivity factors."""
    kcell = np.exp(mt_normal(seed, n ** d, sigma)).reshape((n,) * d)
    return _assemble(n, d, _harmonic_faces(kcell, d, axis_scale))


def convdiff_upwind(n: int, d: int, velocity=(1.0, 0.5, 0.25), peclet: float = 10.0) -> sp.csc_matrix:
    """-eps Lap u + b . grad u, first-order upwind, h = 1, eps = |b|_inf h / Pe (nonsymmetric)."""
    b = np.asarray(velocity[:d], dtype=np.float64)
    eps = float(np.max(np.abs(b))) / peclet
    ones = np.full((n,) * d, eps)
    first = lambda ax: ones[tuple([slice(0, 1) if a == ax else slice(None) for a in range(d)])]
    return _assemble(n, d, lambda ax: (ones, first(ax)), upwind=tuple(b))


def column_scale(A: sp.csc_matrix) -> sp.csc_matrix:
    """A <- A D 10 with D_jj = 1/||A(:,j)||_2 (pspaQR.cpp:132-147)."""
    A = A.tocsc(copy=True)
    sq = np.zeros(A.shape[1])
    for j in range(A.shape[1]):
        v = A.data[A.indptr[j]:A.indptr[j + 1]]
        sq[j] = np.sum(v * v)
    # the reference sums in index order; np.sum on tiny columns is sequential enough, but do it
    # explicitly to keep the same rounding
    d = 1.0 / np.sqrt(sq)
    A.data = A.data * np.repeat(d, np.diff(A.indptr)) * 10.0
    return A


@dataclass
class Problem:
    name: str
    A: sp.csc_matrix          # scaled, natural ordering
    b: np.ndarray             # natural ordering
    n_side: int
    dim: int
    nlevels: int
    tol: float
    skip: int = 0
    meta: dict = field(default_factory=dict)
    X: np.ndarray | None = None      # coordinates (d x n) read from a file, in place of the n^d grid
    parts: np.ndarray | None = None  # leaf parts (METIS-style), when there are no coordinates


def load_problem(matrix: str, nlevels: int, tol: float, coordinates: str | None = None, n: int = -1, d: int = -1,
                 skip: int = 0, rhs: str | None = None, seed: int = 0) -> Problem:
    """The reference driver's input path (pspaQR.cpp:85-167): A from a Matrix Market file, scaled
    A <- A D 10; coordinates from an array file (d x N), or the n^d tensor grid when n and d are
    given (they take precedence, as -n/-d overwrite --coordinates), or none -- the tree then orders
    by graph bisection (always for a rectangular A, which is solved in the least-squares sense).
    b (one entry per row) from an array file, else uniform(-1, 1) from std::mt19937(seed)."""
    from . import mmio

    A = mmio.read_sparse(matrix)
    M, N = A.shape
    if M < N:
        raise ValueError("A has fewer rows than columns (%d x %d)" % A.shape)
    X = None
    if M > N and (coordinates or (n > 0 and d > 0)):
        raise ValueError("a rectangular system (least squares) is ordered without coordinates")
    if n > 0 and d > 0:
        if n ** d != N:
            raise ValueError("n^d = %d but A is %d x %d" % (n ** d, N, N))
    else:
        n, d = 0, 0
        if coordinates:
            X = mmio.read_dense(coordinates)
            if X.shape[1] != N:
                raise ValueError("the coordinate file must hold a d x N matrix (it is %d x %d)" % X.shape)
    b = mmio.read_vector(rhs) if rhs else mt_uniform(seed, M)
    name = os.path.splitext(os.path.basename(str(matrix)))[0]
    return Problem(name, column_scale(A), b, n, d, nlevels, tol, skip, X=X)


@dataclass
class Batch:
    """Independent systems factored together: A is their block-diagonal matrix (system i owns
    unknowns offsets[i] .. offsets[i] + sizes[i] - 1, each in its natural grid ordering), X the
    coordinates of all unknowns (d x n), b the right-hand sides stacked the same way."""
    name: str
    systems: list
    A: sp.csc_matrix
    b: np.ndarray
    X: np.ndarray
    sizes: list
    levels: list
    tol: float
    meta: dict = field(default_factory=dict)

    @property
    def offsets(self):
        return np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(int)


def _coords(n: int, d: int) -> np.ndarray:
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64)] * d, indexing="ij")
    return np.stack([g.ravel() for g in grids], axis=0)


BATCH_SIDES = (16, 24, 32, 40, 48)


def make_batch(count: int = 48, sides=BATCH_SIDES, seed0: int = 3, tol: float = 1e-2) -> Batch:
    """The batch workload: ``count`` 3D 7-point diffusion systems, system i with seed seed0 + i.
    From std::mt19937(seed) the first canonical draw picks the grid side uniformly from ``sides``,
    the second s ~ U[0.5, 2]; the following draws give the cell coefficients exp(N(0, s^2))
    (Box-Muller, harmonic face averages as in config 2); the right-hand side is U[-1, 1] from
    seed + 1000.  Levels by the driver's default rule (pspaQR.cpp:108-113)."""
    systems, blocks, bs, Xs, sizes, levels = [], [], [], [], [], []
    for i in range(count):
        seed = seed0 + i
        n = sides[min(len(sides) - 1, int(mt_canonical(seed, 1)[0] * len(sides)))]
        sig = 0.5 + 1.5 * mt_canonical(seed, 2)[1]
        N = n ** 3
        u = mt_canonical(seed, 2 + 2 * ((N + 1) // 2))[2:]
        u1, u2 = u[0::2], u[1::2]
        rad = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.empty(2 * len(u1))
        z[0::2] = rad * np.cos(2.0 * math.pi * u2)
        z[1::2] = rad * np.sin(2.0 * math.pi * u2)
        kcell = np.exp(sig * z[:N]).reshape((n,) * 3)
        A = column_scale(_assemble(n, 3, _harmonic_faces(kcell, 3)))
        b = mt_uniform(seed + 1000, N)
        lv = default_levels(N)
        systems.append(Problem("diffusion3d_%d_s%d" % (n, seed), A, b, n, 3, lv, tol, meta={"sigma": sig}))
        blocks.append(A)
        bs.append(b)
        Xs.append(_coords(n, 3))
        sizes.append(N)
        levels.append(lv)
    A = sp.block_diag(blocks, format="csc")
    A.sort_indices()
    return Batch("batch%d_diffusion3d" % count, systems, A, np.concatenate(bs), np.asfortranarray(np.concatenate(Xs, axis=1)),
                 sizes, levels, tol)


def gradient_rows(n: int, d: int) -> sp.csc_matrix:
    """Forward differences x(i + e_ax) - x(i) on every interior face of the n^d grid, axis by
    axis, faces in natural order: (d n^(d-1) (n-1)) x n^d."""
    N = n ** d
    idx = np.arange(N).reshape((n,) * d)
    rows, cols, vals = [], [], []
    r0 = 0
    for ax in range(d):
        lo = idx.take(np.arange(n - 1), axis=ax).ravel()
        hi = idx.take(np.arange(1, n), axis=ax).ravel()
        k = np.arange(r0, r0 + lo.size)
        rows += [k, k]
        cols += [hi, lo]
        vals += [np.ones(lo.size), -np.ones(lo.size)]
        r0 += lo.size
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(r0, N))


def least_squares(n: int, d: int, seed: int = 11, lam: float = 0.5, tol: float = 1e-2, nlevels: int | None = None) -> Problem:
    """An overdetermined system A = [K; lam G] (m > n): K the variable-coefficient diffusion
    operator of config 2's kind on the n^d grid (coefficients exp(N(0,1)) from seed), G the
    gradient rows (gradient_rows), then the driver's column scaling A <- A D 10.  b ~ U[-1, 1]
    (m entries) from std::mt19937(seed + 1).  Solved in the least-squares sense."""
    K = diffusion_varcoef(n, d, seed=seed, sigma=1.0)
    A = sp.vstack([K, lam * gradient_rows(n, d)], format="csc")
    A.sort_indices()
    m, N = A.shape
    lv = nlevels if nlevels is not None else max(2, default_levels(N))
    return Problem("ls_diffusion%dd_%d" % (d, n), column_scale(A), mt_uniform(seed + 1, m), n, d, lv, tol,
                   meta={"m": m})


def default_levels(ncols: int) -> int:
    """nlevels = ceil(log2(ncols/64)) with integer division (pspaQR.cpp:108-113)."""
    return int(math.ceil(math.log2(ncols // 64)))


def make_problem(cfg: int | str, scale_side: int | None = None) -> Problem:
    """The five workloads; ``scale_side`` shrinks the grid (parity tests at oracle sizes)."""
    if cfg in (1, "poisson2d"):
        n = scale_side or 64
        A = poisson(n, 2)
        b = mt_uniform(0, n * n)
        lv = 6 if scale_side is None else max(2, default_levels(n * n))
        return Problem("poisson2d_%d" % n, column_scale(A), b, n, 2, lv, 1e-2)
    if cfg in (2, "diffusion3d"):
        n = scale_side or 32
        A = diffusion_varcoef(n, 3, seed=1, sigma=1.0)
        b = mt_uniform(1, n ** 3)
        lv = 9 if scale_side is None else max(2, default_levels(n ** 3))
        return Problem("diffusion3d_%d" % n, column_scale(A), b, n, 3, lv, 1e-3)
    if cfg in (3, "convdiff3d"):
        n = scale_side or 64
        A = convdiff_upwind(n, 3)
        b = mt_normal(2, n ** 3)
        lv = 12 if scale_side is None else max(2, default_levels(n ** 3))
        return Problem("convdiff3d_%d" % n, column_scale(A), b, n, 3, lv, 1e-2, meta={"restart": 50})
    if cfg in (5, "anisotropic3d"):
        n = scale_side or 96
        A = diffusion_varcoef(n, 3, seed=7, sigma=0.5, axis_scale=(1.0, 1e-2, 1e-4))
        b = mt_normal(7 + 1000, n ** 3)
        lv = default_levels(n ** 3)
        return Problem("anisotropic3d_%d" % n, column_scale(A), b, n, 3, lv, 1e-2)
    if cfg in (4, "batch"):
        raise ValueError("config 4 is a batch of systems: use make_batch()")
    raise ValueError("unknown config %r" % (cfg,))
