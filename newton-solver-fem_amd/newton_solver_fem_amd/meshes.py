"""Structured Tet4/Hex8 lattices and the benchmark scene definitions.

The Tet4 lattice reproduces the node numbering and the five-tetrahedron
parity split of the reference's ``examples/assets/tet_grid.usda`` fixture
(2x2x2 cells, 27 nodes, 40 tets): node ``(ix, iy, iz)`` has id
``ix*(ny+1)*(nz+1) + iy*(nz+1) + iz``, cells are enumerated with z fastest,
and a cell's split depends on ``(ix+iy+iz) % 2``.  The Hex8 lattice follows
``examples/assets/hex_grid.usda`` (VTK corner order).

Masses are lumped the way the reference's test fixture does it
(``tests/fem_fixture.py:24-31``): each tetrahedron gives ``rho*V/4`` to each
corner; each hexahedron gives ``rho * sum_g w_g N_a(g)``, which is
``rho*V/8`` per corner for a parallelepiped cell.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Local corner offsets of the five tetrahedra in an even / odd parity cell.
_EVEN = (
    ((1, 0, 0), (1, 0, 1), (1, 1, 0), (0, 0, 0)),
    ((0, 0, 1), (0, 0, 0), (0, 1, 1), (1, 0, 1)),
    ((0, 1, 0), (0, 1, 1), (0, 0, 0), (1, 1, 0)),
    ((1, 1, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1)),
    ((1, 1, 0), (1, 0, 1), (0, 1, 1), (0, 0, 0)),
)
_ODD = (
    ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)),
    ((1, 0, 1), (0, 0, 1), (1, 1, 1), (1, 0, 0)),
    ((1, 1, 0), (0, 1, 0), (1, 0, 0), (1, 1, 1)),
    ((0, 1, 1), (1, 1, 1), (0, 0, 1), (0, 1, 0)),
    ((0, 1, 0), (1, 0, 0), (1, 1, 1), (0, 0, 1)),
)
# VTK Hex8 corner order: reference sign table builder.py:104-107.
_HEX = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def lattice_points(nx: int, ny: int, nz: int, spacing: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Node positions of an (nx+1)x(ny+1)x(nz+1) grid, z fastest, float64."""
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64) * spacing
    return pts + np.asarray(origin, dtype=np.float64)


def _node_id(ix, iy, iz, ny, nz):
    return ix * (ny + 1) * (nz + 1) + iy * (nz + 1) + iz


def tet_lattice_indices(nx: int, ny: int, nz: int) -> np.ndarray:
    """(5*nx*ny*nz, 4) int32 connectivity of the parity-split lattice."""
    cx, cy, cz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cx, cy, cz = cx.ravel(), cy.ravel(), cz.ravel()
    odd = ((cx + cy + cz) % 2).astype(bool)
    out = np.empty((cx.size, 5, 4), dtype=np.int64)
    for table, mask in ((_EVEN, ~odd), (_ODD, odd)):
        for t, tet in enumerate(table):
            for c, (dx, dy, dz) in enumerate(tet):
                out[mask, t, c] = _node_id(cx[mask] + dx, cy[mask] + dy, cz[mask] + dz, ny, nz)
    return out.reshape(-1, 4).astype(np.int32)


def hex_lattice_indices(nx: int, ny: int, nz: int) -> np.ndarray:
    """(nx*ny*nz, 8) int32 VTK-ordered connectivity, cells z fastest."""
    cx, cy, cz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cx, cy, cz = cx.ravel(), cy.ravel(), cz.ravel()
    out = np.empty((cx.size, 8), dtype=np.int64)
    for c, (dx, dy, dz) in enumerate(_HEX):
        out[:, c] = _node_id(cx + dx, cy + dy, cz + dz, ny, nz)
    return out.astype(np.int32)


def lame(young: np.ndarray | float, poisson: np.ndarray | float):
    """(mu, lambda) from Young's modulus and Poisson ratio."""
    young = np.asarray(young, dtype=np.float64)
    poisson = np.asarray(poisson, dtype=np.float64)
    mu = young / (2.0 * (1.0 + poisson))
    lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    return mu, lam


def tet_lumped_masses(points: np.ndarray, tets: np.ndarray, density) -> np.ndarray:
    """rho*V/4 per tetrahedron corner (reference tests/fem_fixture.py:26-28)."""
    pts = np.asarray(points, dtype=np.float64)
    edges = pts[tets[:, 1:]] - pts[tets[:, :1]]
    volume = np.linalg.det(edges) / 6.0
    masses = np.zeros(len(pts))
    rho = np.broadcast_to(np.asarray(density, dtype=np.float64), (len(tets),))
    np.add.at(masses, tets.ravel(), np.repeat(rho * volume / 4.0, 4))
    return masses


@dataclass
class TetScene:
    """One solid described by arrays; ``build`` turns it into a model."""

    points: np.ndarray
    tets: np.ndarray
    materials: np.ndarray  # (T, 3) mu, lambda, k_damp
    masses: np.ndarray
    fixed: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    radius: float = 0.0
    activations: np.ndarray | None = None


def tet_drop_scene(n: int = 10, spacing: float = 0.01, height: float = 0.05, young=1.0e5, poisson=0.45,
                   density=1000.0) -> TetScene:
    """n^3 cells x 5 tets, bottom face at ``height`` (Tet4 drop workload)."""
    pts = lattice_points(n, n, n, spacing, origin=(0.0, 0.0, height))
    tets = tet_lattice_indices(n, n, n)
    mu, lam = lame(young, poisson)
    mats = np.tile(np.array([mu, lam, 0.0]), (len(tets), 1))
    return TetScene(pts, tets, mats, tet_lumped_masses(pts, tets, density))


def large_tet_scene(nx: int = 60, ny: int = 60, nz: int = 56, spacing: float = 0.01, seed: int = 0,
                    jitter: float = 0.2, poisson=0.45, density=1000.0, e_range=(5.0e4, 5.0e5)) -> TetScene:
    """Jittered lattice with log-uniform per-element Young's modulus, x=0 face fixed.

    One ``numpy.random.default_rng(seed)`` stream draws first the interior
    node jitter (uniform in +-jitter*spacing per axis, node-id order), then
    one Young's modulus per element.
    """
    rng = np.random.default_rng(seed)
    pts = lattice_points(nx, ny, nz, spacing)
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    interior = ((ix > 0) & (ix < nx) & (iy > 0) & (iy < ny) & (iz > 0) & (iz < nz)).ravel()
    pts[interior] += rng.uniform(-jitter, jitter, size=(int(interior.sum()), 3)) * spacing
    tets = tet_lattice_indices(nx, ny, nz)
    young = np.exp(rng.uniform(math.log(e_range[0]), math.log(e_range[1]), size=len(tets)))
    mu, lam = lame(young, poisson)
    mats = np.stack([mu, lam, np.zeros_like(mu)], axis=1)
    fixed = np.flatnonzero(ix.ravel() == 0).astype(np.int32)
    return TetScene(pts, tets, mats, tet_lumped_masses(pts, tets, density), fixed=fixed)


def tet_bar_scene(nx: int = 40, ny: int = 8, nz: int = 8, spacing: float = 0.01, young=1.0e5, poisson=0.45,
                  density=1000.0, seed: int = 2, fix_base: bool = True) -> TetScene:
    """Cantilever bar with seeded per-element activations in [0, 0.3]."""
    pts = lattice_points(nx, ny, nz, spacing)
    tets = tet_lattice_indices(nx, ny, nz)
    mu, lam = lame(young, poisson)
    mats = np.tile(np.array([mu, lam, 0.0]), (len(tets), 1))
    rng = np.random.default_rng(seed)
    act = rng.uniform(0.0, 0.3, size=len(tets))
    fixed = np.flatnonzero(pts[:, 0] == 0.0).astype(np.int32) if fix_base else np.zeros(0, np.int32)
    return TetScene(pts, tets, mats, tet_lumped_masses(pts, tets, density), fixed=fixed, activations=act)


@dataclass
class HexScene:
    points: np.ndarray
    hexes: np.ndarray
    materials: np.ndarray  # (H, 3)
    density: float
    fixed: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    radius: float = 0.0


def hex_cantilever_scene(nx: int = 64, ny: int = 16, nz: int = 16, spacing: float = 0.005, young=1.0e6,
                         poisson=0.4, density=1100.0) -> HexScene:
    """Hex8 cantilever with the x=0 face fixed."""
    pts = lattice_points(nx, ny, nz, spacing)
    hexes = hex_lattice_indices(nx, ny, nz)
    mu, lam = lame(young, poisson)
    mats = np.tile(np.array([mu, lam, 0.0]), (len(hexes), 1))
    fixed = np.flatnonzero(pts[:, 0] == 0.0).astype(np.int32)
    return HexScene(pts, hexes, mats, density, fixed=fixed)


def batched_tet_scenes(count: int = 256, seed: int = 1, t_range=(500, 50_000), spacing: float = 0.01):
    """Uneven independent Tet4 lattices for the batched workload.

    Per instance, in this order from one ``default_rng(seed)`` stream: a
    target tet count log-uniform in ``t_range``, an aspect ratio uniform in
    [1, 4], Young's modulus uniform in [1e4, 1e6], Poisson ratio uniform in
    [0.3, 0.48] and a drop height uniform in [0, 0.1] m.  Cell counts are
    ``(round(a*n), n, n)`` with ``5*a*n^3`` closest to the target.
    """
    rng = np.random.default_rng(seed)
    scenes = []
    for _ in range(count):
        target = math.exp(rng.uniform(math.log(t_range[0]), math.log(t_range[1])))
        aspect = rng.uniform(1.0, 4.0)
        young = rng.uniform(1.0e4, 1.0e6)
        poisson = rng.uniform(0.3, 0.48)
        height = rng.uniform(0.0, 0.1)
        n = max(1, int(round((target / (5.0 * aspect)) ** (1.0 / 3.0))))
        nx = max(1, int(round(aspect * n)))
        pts = lattice_points(nx, n, n, spacing, origin=(0.0, 0.0, height))
        tets = tet_lattice_indices(nx, n, n)
        mu, lam = lame(young, poisson)
        mats = np.tile(np.array([float(mu), float(lam), 0.0]), (len(tets), 1))
        scenes.append(TetScene(pts, tets, mats, tet_lumped_masses(pts, tets, 1000.0)))
    return scenes
