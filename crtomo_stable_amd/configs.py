"""Seeded synthetic workloads of the forward / Jacobian / CG path.

* ``c0``: CRMod 2.5D forward, 60x20 quads at 0.5 m, 30 surface electrodes at
  1 m, dipole-dipole a=1, n=1..8 (188 quadrupoles), 100 ohm-m with a 4x2 m
  block of 10 ohm-m at 2 m depth.
* ``c1``: complex forward + sensitivities on a seeded Delaunay mesh of ~20k
  triangles over 100x30 m with 2 m sinusoidal topography, 64 electrodes at
  1.5 m, 2000 dipole-dipole + Wenner quadrupoles, Gaussian-random-field
  magnitude 10..1000 ohm-m (5 m correlation) and phase -1..-30 mrad.
* ``c2_model(i)``: model i of the batched set, a seeded quad mesh of 5k..40k
  cells at aspect 4:1 with 48..96 surface electrodes and a Gaussian-random-field
  model (2..20 m correlation, 10..1000 ohm-m, -1..-30 mrad).
* ``c3``: 400x100 quads, 96 electrodes at 2 m, 4500 dipole-dipole/Wenner
  quadrupoles, three blocks in a layered 50/200 ohm-m background.
* ``small``: reduced versions of the same generators for parity tests.

Grids use the reference fixture conventions (type 12 surface, type 11 sides
and bottom) and are Cuthill-McKee renumbered by the library's CutMcK.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import grid as G


@dataclass
class Workload:
    name: str
    grid: G.Grid
    survey: G.Survey
    sigma: np.ndarray
    notes: str = ""


def _cutmck(grid: G.Grid, cutmck: Optional[Callable] = None) -> G.Grid:
    if cutmck is None:
        from .lib import cutmck as lib_cutmck
        cutmck = lib_cutmck
    perm, mb = cutmck(grid)[:2]
    return G.renumber(grid, perm, mb)


def quad_workload(name: str, nx: int, nz: int, dx: float, dz: float, nelec: int, elec_every: int,
                  elec_first: int, quads, sigma_fn, cutmck=None, mixed: bool = True,
                  dirichlet: bool = False) -> Workload:
    n1 = nx + 1
    elec_nodes = [elec_first + i * elec_every + 1 for i in range(nelec)]
    g = G.quad_grid(nx, nz, dx, dz, elec_nodes=elec_nodes, mixed=mixed, dirichlet=dirichlet)
    cx, cy = g.centroids()
    sigma = sigma_fn(cx, cy)
    g = _cutmck(g, cutmck)
    assert n1 >= elec_nodes[-1]
    return Workload(name, g, G.encode(quads), sigma)


def c0(cutmck=None) -> Workload:
    def sig(cx, cy):
        rho = np.full(cx.size, 100.0)
        blk = (cx > 13.0) & (cx < 17.0) & (cy < -2.0) & (cy > -4.0)
        rho[blk] = 10.0
        return G.sigma_from_rho(rho, np.zeros_like(rho))
    quads = G.dipole_dipole(30, [1], range(1, 9))
    assert len(quads) == 188
    return quad_workload("c0", 60, 20, 0.5, 0.5, 30, 2, 1, quads, sig, cutmck)


def c1(cutmck=None, scale: float = 1.0, nmeas: int = 2000, nelec: int = 64, seed: int = 1) -> Workload:
    """``scale`` < 1 shrinks the domain (and electrode count) for parity tests."""
    length, depth = 100.0 * scale, 30.0 * scale
    spacing = 1.5
    span = spacing * (nelec - 1)
    ex = (length - span) / 2.0 + spacing * np.arange(nelec)
    h = 0.55
    g = G.delaunay_mesh(length, depth, h, ex, topo_amp=2.0 * scale, topo_wavelength=50.0 * scale, seed=seed)
    cx, cy = g.centroids()
    f = G.uniformize(G.gaussian_random_field(cx, cy, 5.0, seed=seed + 100))
    mag = 10.0 ** (1.0 + 2.0 * f)
    fp = G.uniformize(G.gaussian_random_field(cx, cy, 10.0, seed=seed + 200))
    pha = -1.0 - 29.0 * fp
    sigma = G.sigma_from_rho(mag, pha)
    quads = G.dipole_dipole(nelec, [1, 2, 3, 4], range(1, 9)) + G.wenner(nelec, range(1, nelec // 3 + 1))
    quads = quads[:nmeas]
    g = _cutmck(g, cutmck)
    return Workload("c1", g, G.encode(quads), sigma,
                    notes=f"{g.elanz} triangles, {g.sanz} nodes, mb={g.mb}")


def c2_model(i: int, cutmck=None, seed: int = 2, min_cells: int = 5000, max_cells: int = 40000) -> Workload:
    """Model ``i`` of the batched forward set: cell count uniform in
    [min_cells, max_cells] at nx = 4 nz (0.5 m cells), 48..96 electrodes spread over
    the surface, dipole-dipole a = 1..2, n = 1..8, magnitude 10..1000 ohm-m and
    phase -1..-30 mrad as Gaussian random fields of one correlation length in
    2..20 m.  Everything follows from (seed, i)."""
    rng = np.random.default_rng([seed, i])
    cells = rng.uniform(min_cells, max_cells)
    nz = max(4, int(round(np.sqrt(cells / 4.0))))
    nx = 4 * nz
    nelec = int(min(rng.integers(48, 97), nx // 2))
    spacing_nodes = max(1, (nx - 4) // (nelec - 1))
    first = (nx - spacing_nodes * (nelec - 1)) // 2
    corr = float(rng.uniform(2.0, 20.0))
    fseed = int(rng.integers(1 << 30))

    def sig(cx, cy):
        f = G.uniformize(G.gaussian_random_field(cx, cy, corr, seed=fseed))
        fp = G.uniformize(G.gaussian_random_field(cx, cy, corr, seed=fseed + 1))
        return G.sigma_from_rho(10.0 ** (1.0 + 2.0 * f), -1.0 - 29.0 * fp)

    quads = G.dipole_dipole(nelec, [1, 2], range(1, 9))
    return quad_workload(f"c2[{i}]", nx, nz, 0.5, 0.5, nelec, spacing_nodes, first, quads, sig, cutmck)


def grf_sigma(grid: G.Grid, corr: float, seed: int) -> np.ndarray:
    """A complex conductivity on ``grid``: magnitude log-uniform 10..1000 ohm-m and phase
    -1..-30 mrad, each a Gaussian random field of correlation length ``corr`` (m)."""
    cx, cy = grid.centroids()
    f = G.uniformize(G.gaussian_random_field(cx, cy, corr, seed=seed))
    fp = G.uniformize(G.gaussian_random_field(cx, cy, corr, seed=seed + 1))
    return G.sigma_from_rho(10.0 ** (1.0 + 2.0 * f), -1.0 - 29.0 * fp)


def c3(cutmck=None, nx: int = 400, nz: int = 100, nelec: int = 96, nmeas: int = 4500, seed: int = 3) -> Workload:
    dx = dz = 1.0 * (400 / nx) * 0.5 if nx != 400 else 0.5
    spacing_nodes = max(1, (nx - 4) // (nelec - 1))
    first = (nx - spacing_nodes * (nelec - 1)) // 2
    rng = np.random.default_rng(seed)

    def sig(cx, cy):
        rho = np.where(cy > -0.25 * nz * dz, 50.0, 200.0)
        for _ in range(3):
            w = rng.uniform(0.08, 0.2) * nx * dx
            hgt = rng.uniform(0.1, 0.25) * nz * dz
            x0 = rng.uniform(0.1, 0.9 - w / (nx * dx)) * nx * dx
            z0 = -rng.uniform(0.05, 0.6) * nz * dz
            val = rng.choice([10.0, 1000.0])
            blk = (cx > x0) & (cx < x0 + w) & (cy < z0) & (cy > z0 - hgt)
            rho[blk] = val
        return G.sigma_from_rho(rho, np.full(rho.size, -10.0))

    quads = G.dipole_dipole(nelec, [1, 2, 3, 4, 5, 6], range(1, 9)) + G.wenner(nelec, range(1, nelec // 3 + 1))
    if len(quads) > nmeas:
        idx = np.sort(np.random.default_rng(seed + 1).choice(len(quads), nmeas, replace=False))
        quads = [quads[i] for i in idx]
    return quad_workload("c3", nx, nz, dx, dz, nelec, spacing_nodes, first, quads, sig, cutmck)


def small_quad(nx: int = 24, nz: int = 8, nelec: int = 10, cutmck=None, mixed: bool = True,
               dirichlet: bool = False) -> Workload:
    """A small C0-like grid for fast parity tests (type-8 quads; mixed boundaries
    unless ``mixed=False``, which leaves every boundary no-flow, or ``dirichlet=True``,
    which makes the sides and the bottom Dirichlet boundaries, type 13)."""
    def sig(cx, cy):
        rho = 30.0 + 70.0 * (cx / cx.max())
        blk = (cx > 0.3 * cx.max()) & (cx < 0.5 * cx.max()) & (cy < -1.0)
        rho[blk] = 5.0
        return G.sigma_from_rho(rho, -5.0 - 10.0 * (cx / cx.max()))
    quads = G.dipole_dipole(nelec, [1, 2], range(1, 5))
    name = "small_quad_dirichlet" if dirichlet else ("small_quad" if mixed else "small_quad_noflow")
    return quad_workload(name, nx, nz, 0.5, 0.5, nelec, 2, 2, quads, sig, cutmck, mixed, dirichlet)
