"""Grid, electrode and measurement data of the CRTomo forward path.

The arrays keep the meaning, numbering and layout of the reference's Fortran
module data (``src/elemmod.f90``, ``src/electrmod.f90``, ``src/datmod.f90``):

* node, element, electrode and model numbers are 1-based;
* ``nrel`` is ``(neltot, maxn)`` (element-major), area elements first, then the
  boundary elements whose number ``rel`` is ``iel - elanz`` (``src/precal.f90:121``);
* ``x``/``y`` hold the coordinates of each node (``sx(snr(j))``, ``sy(snr(j))``);
* measurements are encoded as ``strnr = e2*10000 + e1`` (current, ``+I`` at e2,
  ``-I`` at e1) and ``vnr = e4*10000 + e3`` (potential ``phi(e4) - phi(e3)``),
  ``src/bvolt.f90:44-63``.

Readers/writers follow the text formats of ``src/relem.f90:60-151``,
``src/relectr.f90``, ``src/rdatm.f90`` and ``cutmckee/welem.f:40-70``.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence, Tuple

import numpy as np

AREA_TYPES = (3, 5, 8)
BOUNDARY_TYPES = (11, 12, 13)


@dataclass
class Grid:
    sanz: int
    mb: int
    typ: np.ndarray        # (typanz,) element type per group
    nelanz: np.ndarray     # (typanz,) elements per group
    selanz: np.ndarray     # (typanz,) nodes per element per group
    nrel: np.ndarray       # (neltot, maxn) int32, 1-based node numbers (0 padded)
    x: np.ndarray          # (sanz,) float64
    y: np.ndarray          # (sanz,) float64
    rnr: np.ndarray        # (relanz,) int32, 1-based area element next to each boundary element
    enr: np.ndarray        # (eanz,) int32, 1-based electrode nodes
    nx: int = 0            # regular-grid cell counts (0 when unstructured)
    nz: int = 0

    def __post_init__(self):
        self.typ = np.ascontiguousarray(self.typ, dtype=np.int32)
        self.nelanz = np.ascontiguousarray(self.nelanz, dtype=np.int32)
        self.selanz = np.ascontiguousarray(self.selanz, dtype=np.int32)
        self.nrel = np.ascontiguousarray(self.nrel, dtype=np.int32)
        self.x = np.ascontiguousarray(self.x, dtype=np.float64)
        self.y = np.ascontiguousarray(self.y, dtype=np.float64)
        self.rnr = np.ascontiguousarray(self.rnr, dtype=np.int32)
        self.enr = np.ascontiguousarray(self.enr, dtype=np.int32)

    @property
    def typanz(self) -> int:
        return int(self.typ.size)

    @property
    def maxn(self) -> int:
        return int(self.selanz.max())

    @property
    def elanz(self) -> int:
        return int(self.nelanz[self.typ < 10].sum())

    @property
    def relanz(self) -> int:
        return int(self.nelanz[self.typ > 10].sum())

    @property
    def eanz(self) -> int:
        return int(self.enr.size)

    @property
    def npk(self) -> int:
        return self.maxn * (self.maxn + 1) // 2

    def area_nodes(self) -> np.ndarray:
        """(elanz, maxn) node numbers of the area elements."""
        return self.nrel[: self.elanz]

    def centroids(self) -> Tuple[np.ndarray, np.ndarray]:
        """Element centres as ``src/relem.f90:128-138``."""
        out_x = np.zeros(self.elanz)
        out_y = np.zeros(self.elanz)
        off = 0
        for t, n, s in zip(self.typ, self.nelanz, self.selanz):
            if t < 10:
                idx = self.nrel[off:off + n, :s] - 1
                out_x[off:off + n] = self.x[idx].sum(axis=1) / s
                out_y[off:off + n] = self.y[idx].sum(axis=1) / s
            off += n
        return out_x, out_y

    def validate(self) -> None:
        """Band and ordering assumptions of the path (``src/kompab.f90:98-102``)."""
        groups = list(self.typ)
        area_seen_after_boundary = False
        seen_boundary = False
        for t in groups:
            if t > 10:
                seen_boundary = True
            elif seen_boundary:
                area_seen_after_boundary = True
        if area_seen_after_boundary:
            raise ValueError("area element groups must precede boundary groups (src/precal.f90:121)")
        off = 0
        for t, n, s in zip(self.typ, self.nelanz, self.selanz):
            if t != 12:
                nodes = self.nrel[off:off + n, :s]
                bw = int((nodes.max(axis=1) - nodes.min(axis=1)).max()) if n else 0
                if bw > self.mb:
                    raise ValueError(f"element band width {bw} exceeds mb={self.mb} (errnr 19)")
            off += n


@dataclass
class Survey:
    strnr: np.ndarray  # (nanz,) int32
    vnr: np.ndarray    # (nanz,) int32

    def __post_init__(self):
        self.strnr = np.ascontiguousarray(self.strnr, dtype=np.int32)
        self.vnr = np.ascontiguousarray(self.vnr, dtype=np.int32)

    @property
    def nanz(self) -> int:
        return int(self.strnr.size)

    def electrodes(self) -> np.ndarray:
        """(nanz, 4) electrode numbers (e1, e2, e3, e4), ``src/bvolt.f90:44-52``."""
        e1 = self.strnr % 10000
        e2 = (self.strnr - e1) // 10000
        e3 = self.vnr % 10000
        e4 = (self.vnr - e3) // 10000
        return np.stack([e1, e2, e3, e4], axis=1).astype(np.int32)


def sigma_from_rho(mag: np.ndarray, pha_mrad: np.ndarray) -> np.ndarray:
    """Complex conductivity from resistivity magnitude/phase (``src/rsigma.f90:168``)."""
    pha = np.asarray(pha_mrad, dtype=np.float64) / 1e3
    bet = np.asarray(mag, dtype=np.float64)
    return (np.cos(pha) / bet - 1j * np.sin(pha) / bet).astype(np.complex128)


# ----------------------------------------------------------------- file formats
def _tokens(path: str) -> List[str]:
    with open(path) as f:
        return f.read().split()


def read_elem(path: str) -> Grid:
    """Read an ``elem.dat`` (``src/relem.f90:60-151``)."""
    tok = _tokens(path)
    p = 0

    def nxt(conv=int):
        nonlocal p
        v = conv(tok[p])
        p += 1
        return v

    sanz, typanz, mb = nxt(), nxt(), nxt()
    typ, nel, sel = [], [], []
    for _ in range(typanz):
        typ.append(nxt()); nel.append(nxt()); sel.append(nxt())
    snr = np.zeros(sanz, np.int32)
    sx = np.zeros(sanz)
    sy = np.zeros(sanz)
    for i in range(sanz):
        snr[i] = nxt(); sx[i] = nxt(float); sy[i] = nxt(float)
    maxn = max(sel)
    neltot = sum(nel)
    nrel = np.zeros((neltot, maxn), np.int32)
    iel = 0
    for t, n, s in zip(typ, nel, sel):
        for _ in range(n):
            for k in range(s):
                nrel[iel, k] = nxt()
            iel += 1
    relanz = sum(n for t, n in zip(typ, nel) if t > 10)
    rnr = np.array([nxt() for _ in range(relanz)], np.int32)
    x = sx[snr - 1]
    y = sy[snr - 1]
    return Grid(sanz=sanz, mb=mb, typ=np.array(typ), nelanz=np.array(nel), selanz=np.array(sel),
                nrel=nrel, x=x, y=y, rnr=rnr, enr=np.zeros(0, np.int32))


def fix_rnr(grid: Grid) -> Grid:
    """Border-to-element check of ``src/relem.f90:197-229``: every boundary element
    is attached to the area element (of the first type group) holding both its
    nodes."""
    elanz = grid.elanz
    s0 = int(grid.selanz[0])
    owner = {}
    for k in range(elanz):
        nodes = grid.nrel[k, :s0]
        for a in range(s0):
            for b in range(a + 1, s0):
                owner[(min(nodes[a], nodes[b]), max(nodes[a], nodes[b]))] = k + 1
    rnr = grid.rnr.copy()
    for i in range(grid.relanz):
        n1, n2 = grid.nrel[elanz + i, 0], grid.nrel[elanz + i, 1]
        k = owner.get((min(n1, n2), max(n1, n2)))
        if k is not None and rnr[i] != k:
            rnr[i] = k
    return replace(grid, rnr=rnr)


def read_rho(path: str) -> np.ndarray:
    """``rho.dat``: elanz, then magnitude (ohm m) and phase (mrad) per element;
    returns sigma (``src/rsigma.f90:168``)."""
    tok = _tokens(path)
    n = int(tok[0])
    v = np.array(tok[1:1 + 2 * n], dtype=float).reshape(n, 2)
    return sigma_from_rho(v[:, 0], v[:, 1])


def read_elem_records(path: str):
    """Raw records of an ``elem.dat``: header, type table, (snr, sx, sy), element lists, rnr."""
    tok = _tokens(path)
    sanz, typanz, mb = int(tok[0]), int(tok[1]), int(tok[2])
    p = 3
    types = []
    for _ in range(typanz):
        types.append(tuple(int(t) for t in tok[p:p + 3]))
        p += 3
    rec = np.array(tok[p:p + 3 * sanz], dtype=object).reshape(sanz, 3)
    p += 3 * sanz
    return sanz, typanz, mb, types, rec, tok[p:]


def read_elec(path: str) -> np.ndarray:
    tok = _tokens(path)
    n = int(tok[0])
    return np.array([int(t) for t in tok[1:1 + n]], np.int32)


def read_config(path: str) -> Survey:
    tok = _tokens(path)
    n = int(tok[0])
    vals = np.array(tok[1:1 + 2 * n], dtype=np.int64).reshape(n, 2)
    return Survey(strnr=vals[:, 0], vnr=vals[:, 1])


def read_volt(path: str):
    """``volt.dat``: nanz, then strnr vnr magnitude phase(mrad)."""
    tok = _tokens(path)
    n = int(tok[0])
    rows = np.array(tok[1:1 + 4 * n], dtype=object).reshape(n, 4)
    survey = Survey(strnr=rows[:, 0].astype(np.int64), vnr=rows[:, 1].astype(np.int64))
    return survey, rows[:, 2].astype(float), rows[:, 3].astype(float)


# ----------------------------------------------------------- synthetic grids
def quad_grid(nx: int, nz: int, dx: float, dz: float, x0: float = 0.0,
              elec_nodes: Optional[Sequence[int]] = None, mixed: bool = True, dirichlet: bool = False) -> Grid:
    """Regular quad grid in the conventions of the reference fixtures
    (``tests/minimal/Reference/grid/elem.dat.orig``): nodes row-major from the
    surface, type-8 quads counter-clockwise from bottom-left, type-12 no-flow
    surface left to right, type-11 mixed boundary clockwise (right side down,
    bottom right to left, left side up).  With ``mixed=False`` the sides and
    the bottom are no-flow (type 12) as well, so no system depends on the
    source electrode (lbeta false, one factorisation per wavenumber).  With
    ``dirichlet=True`` the sides and the bottom are type-13 elements (zero potential,
    src/randb.f90) instead.  Not yet CutMcK-ordered (mb = 0)."""
    mixed_flag = mixed
    n1 = nx + 1
    ix, iz = np.meshgrid(np.arange(n1), np.arange(nz + 1))
    x = (x0 + ix * dx).ravel().astype(np.float64)
    y = (-iz * dz).ravel().astype(np.float64)

    def node(i, k):
        return k * n1 + i + 1

    cells = []
    for k in range(nz):
        for i in range(nx):
            cells.append((node(i, k + 1), node(i + 1, k + 1), node(i + 1, k), node(i, k)))
    surf = [(node(i, 0), node(i + 1, 0)) for i in range(nx)]
    # rnr covers every boundary element in file order (surface group, then mixed)
    rnr = [i + 1 for i in range(nx)]
    mixed = []
    for k in range(nz):                       # right side, top -> bottom
        mixed.append((node(nx, k), node(nx, k + 1))); rnr.append(k * nx + nx)
    for i in range(nx, 0, -1):                # bottom, right -> left
        mixed.append((node(i, nz), node(i - 1, nz))); rnr.append((nz - 1) * nx + i)
    for k in range(nz, 0, -1):                # left side, bottom -> top
        mixed.append((node(0, k), node(0, k - 1))); rnr.append((k - 1) * nx + 1)
    neltot = len(cells) + len(surf) + len(mixed)
    nrel = np.zeros((neltot, 4), np.int32)
    nrel[: len(cells)] = np.array(cells)
    nrel[len(cells): len(cells) + len(surf), :2] = np.array(surf)
    nrel[len(cells) + len(surf):, :2] = np.array(mixed)
    enr = np.array([] if elec_nodes is None else elec_nodes, np.int32)
    if dirichlet:
        typ, nelanz, selanz = [8, 12, 13], [len(cells), len(surf), len(mixed)], [4, 2, 2]
    elif mixed_flag:
        typ, nelanz, selanz = [8, 12, 11], [len(cells), len(surf), len(mixed)], [4, 2, 2]
    else:
        typ, nelanz, selanz = [8, 12], [len(cells), len(surf) + len(mixed)], [4, 2]
    return Grid(sanz=n1 * (nz + 1), mb=0, typ=np.array(typ), nelanz=np.array(nelanz),
                selanz=np.array(selanz), nrel=nrel, x=x, y=y,
                rnr=np.array(rnr, np.int32), enr=enr, nx=nx, nz=nz)


def delaunay_mesh(length: float, depth: float, h: float, elec_x: Sequence[float],
                  topo_amp: float, topo_wavelength: float, seed: int) -> Grid:
    """Seeded Delaunay triangulation of ``[0,length] x [-depth, top(x)]`` with a
    sinusoidal surface ``top(x) = topo_amp*sin(2*pi*x/topo_wavelength)`` that
    decays linearly to a flat bottom.  Interior points form a jittered lattice
    of spacing ``h``; electrode positions are surface nodes.  Boundary elements
    follow the reference conventions (type 12 surface, type 11 elsewhere,
    clockwise so that ``src/beta.f90:92-93`` gets the outward normal)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    elec_x = np.asarray(elec_x, dtype=np.float64)
    # surface points: electrodes plus fill-in at spacing ~h
    sx = [0.0, length]
    sx += list(elec_x)
    fill = np.arange(0.0, length + 1e-9, h)
    for fx in fill:
        if np.min(np.abs(np.asarray(sx) - fx)) > 0.45 * h:
            sx.append(fx)
    sx = np.unique(np.round(np.array(sx), 9))
    nside = max(2, int(round(depth / h)))
    side_z = -np.linspace(0.0, depth, nside + 1)[1:-1]
    bot_x = np.linspace(0.0, length, max(2, int(round(length / h))) + 1)
    pts = [np.stack([sx, np.zeros_like(sx)], 1),
           np.stack([bot_x, np.full_like(bot_x, -depth)], 1),
           np.stack([np.zeros_like(side_z), side_z], 1),
           np.stack([np.full_like(side_z, length), side_z], 1)]
    gx = np.arange(h, length - 0.5 * h, h)
    gz = -np.arange(h, depth - 0.5 * h, h)
    GX, GZ = np.meshgrid(gx, gz)
    jit = (rng.random((2,) + GX.shape) - 0.5) * 0.5 * h
    inner = np.stack([(GX + jit[0]).ravel(), (GZ + jit[1]).ravel()], 1)
    # keep interior points away from the surface line
    inner = inner[inner[:, 1] < -0.4 * h]
    pts.append(inner)
    P = np.concatenate(pts)
    tri = Delaunay(P, qhull_options="QJ Qbb Qc Qz" if False else "Qbb Qc Qz")
    simp = tri.simplices.astype(np.int64)

    def topo(xv):
        return topo_amp * np.sin(2.0 * np.pi * xv / topo_wavelength)

    X = P[:, 0]
    Z = P[:, 1] + topo(P[:, 0]) * (1.0 + P[:, 1] / depth)
    # counter-clockwise in the mapped coordinates (src/elem3.f90:75: det > 0)
    a = X[simp]
    b = Z[simp]
    det = (a[:, 1] - a[:, 0]) * (b[:, 2] - b[:, 0]) - (a[:, 2] - a[:, 0]) * (b[:, 1] - b[:, 0])
    simp[det < 0] = simp[det < 0][:, [0, 2, 1]]
    # drop degenerate simplices
    a = X[simp]; b = Z[simp]
    det = (a[:, 1] - a[:, 0]) * (b[:, 2] - b[:, 0]) - (a[:, 2] - a[:, 0]) * (b[:, 1] - b[:, 0])
    simp = simp[det > 1e-10]
    # order cells row-major by centroid (surface first) for locality
    cz = Z[simp].mean(1)
    cx = X[simp].mean(1)
    order = np.lexsort((cx, np.round(-cz / h).astype(np.int64)))
    simp = simp[order]
    # boundary edges: edges used by exactly one triangle
    edges = {}
    for t, (p0, p1, p2) in enumerate(simp):
        for u, v in ((p0, p1), (p1, p2), (p2, p0)):
            key = (min(u, v), max(u, v))
            edges.setdefault(key, []).append((t, u, v))
    bnd = [(lst[0][0], lst[0][1], lst[0][2]) for lst in edges.values() if len(lst) == 1]
    # triangles are CCW, so (u -> v) traverses the boundary counter-clockwise;
    # the reference wants clockwise type-11 segments, i.e. (v -> u).
    flat_x, flat_z = P[:, 0], P[:, 1]
    eps = 1e-9
    surf, mixed = [], []
    for t, u, v in bnd:
        if abs(flat_z[u]) < eps and abs(flat_z[v]) < eps:
            surf.append((t, u, v))
        else:
            mixed.append((t, v, u))
    surf.sort(key=lambda e: min(flat_x[e[1]], flat_x[e[2]]))
    surf = [(t, u, v) if flat_x[u] < flat_x[v] else (t, v, u) for t, u, v in surf]
    # chain the mixed boundary clockwise starting at the top-right corner
    nxt = {u: (t, u, v) for t, u, v in mixed}
    start = int(np.argmin(np.where(np.abs(flat_x - length) < eps, flat_z, np.inf) * -1.0 + 0.0))
    corners = np.where((np.abs(flat_x - length) < eps) & (np.abs(flat_z) < eps))[0]
    cur = int(corners[0])
    chain = []
    while cur in nxt and len(chain) < len(mixed):
        e = nxt.pop(cur)
        chain.append(e)
        cur = e[2]
    if len(chain) != len(mixed):
        raise RuntimeError("mixed boundary is not a single chain")
    # renumber nodes: surface first (by x), then the rest row-major
    sanz = P.shape[0]
    key_row = np.round(-flat_z / h).astype(np.int64)
    order_n = np.lexsort((flat_x, key_row))
    newnum = np.empty(sanz, np.int64)
    newnum[order_n] = np.arange(1, sanz + 1)
    cells = newnum[simp]
    surf_e = np.array([(newnum[u], newnum[v]) for _, u, v in surf])
    mix_e = np.array([(newnum[u], newnum[v]) for _, u, v in chain])
    rnr = np.array([t + 1 for t, _, _ in surf] + [t + 1 for t, _, _ in chain], np.int32)
    neltot = len(cells) + len(surf_e) + len(mix_e)
    nrel = np.zeros((neltot, 3), np.int32)
    nrel[: len(cells)] = cells
    nrel[len(cells): len(cells) + len(surf_e), :2] = surf_e
    nrel[len(cells) + len(surf_e):, :2] = mix_e
    xn = np.empty(sanz); yn = np.empty(sanz)
    xn[newnum - 1] = X
    yn[newnum - 1] = Z
    enr = np.array([newnum[np.where((np.abs(flat_x - ex) < 1e-9) & (np.abs(flat_z) < eps))[0][0]]
                    for ex in elec_x], np.int32)
    return Grid(sanz=sanz, mb=0, typ=np.array([3, 12, 11]),
                nelanz=np.array([len(cells), len(surf_e), len(mix_e)]),
                selanz=np.array([3, 2, 2]), nrel=nrel, x=xn, y=yn, rnr=rnr, enr=enr)


def renumber(grid: Grid, perm: np.ndarray, mb: int) -> Grid:
    """Apply a CutMcK permutation (``perm[old-1]`` = new number) to elements,
    node coordinates and electrodes (``cutmckee/cutmck.f:313-337``)."""
    perm = np.asarray(perm, np.int64)
    nrel = grid.nrel.astype(np.int64)
    mask = nrel > 0
    nrel[mask] = perm[nrel[mask] - 1]
    x = np.empty_like(grid.x); y = np.empty_like(grid.y)
    x[perm - 1] = grid.x
    y[perm - 1] = grid.y
    enr = perm[grid.enr.astype(np.int64) - 1] if grid.enr.size else grid.enr
    return replace(grid, mb=int(mb), nrel=nrel.astype(np.int32), x=x, y=y,
                   enr=np.asarray(enr, np.int32))


# ----------------------------------------------------------- synthetic surveys
def dipole_dipole(nelec: int, a_values: Sequence[int], n_values: Sequence[int]) -> List[Tuple[int, int, int, int]]:
    out = []
    for a in a_values:
        for n in n_values:
            for i in range(1, nelec + 1):
                A, B, M, N = i, i + a, i + a + n * a, i + 2 * a + n * a
                if N > nelec:
                    break
                out.append((A, B, M, N))
    return out


def wenner(nelec: int, a_values: Sequence[int]) -> List[Tuple[int, int, int, int]]:
    out = []
    for a in a_values:
        for i in range(1, nelec + 1):
            A, M, N, B = i, i + a, i + 2 * a, i + 3 * a
            if B > nelec:
                break
            out.append((A, B, M, N))
    return out


def encode(quads: Sequence[Tuple[int, int, int, int]]) -> Survey:
    q = np.asarray(quads, np.int64)
    return Survey(strnr=q[:, 0] * 10000 + q[:, 1], vnr=q[:, 2] * 10000 + q[:, 3])


def gaussian_random_field(px: np.ndarray, py: np.ndarray, corr_len: float, seed: int,
                          nmodes: int = 256) -> np.ndarray:
    """Zero-mean unit-variance field with Gaussian covariance of length ``corr_len``
    (random Fourier features), evaluated at the points (px, py)."""
    rng = np.random.default_rng(seed)
    kx = rng.normal(0.0, 1.0 / corr_len, nmodes)
    kz = rng.normal(0.0, 1.0 / corr_len, nmodes)
    ph = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    ux, ix = np.unique(px, return_inverse=True)
    uy, iy = np.unique(py, return_inverse=True)
    if ux.size * uy.size <= 4 * px.size:
        # points on a tensor grid (regular quad meshes): cos(a + b) = cos a cos b - sin a sin b
        # turns the n x nmodes cosines into two small products
        a = np.outer(ux, kx) + ph
        b = np.outer(uy, kz)
        tab = np.cos(a) @ np.cos(b).T - np.sin(a) @ np.sin(b).T
        f = np.sqrt(2.0 / nmodes) * tab[ix, iy]
    else:
        f = np.sqrt(2.0 / nmodes) * np.cos(np.outer(px, kx) + np.outer(py, kz) + ph).sum(1)
    return (f - f.mean()) / max(f.std(), 1e-12)


def uniformize(f: np.ndarray) -> np.ndarray:
    """Map a standard normal field to [0,1] through its CDF."""
    from scipy.special import ndtr
    return ndtr(f)
