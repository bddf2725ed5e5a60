"""Seeded synthetic PFEM meshes: jittered lattices split into positively oriented P1 simplices.

The reference reads gmsh node clouds and triangulates them with CGAL's alpha shape
(``srcs/mesh/Mesh.cpp:762-878``, ``srcs/mesh/Mesh2D.cpp:32``, ``srcs/mesh/Mesh3D.cpp:36``).
Here the connectivity is a Kuhn/Freudenthal split of a lattice instead, which is
reproducible without CGAL and keeps every element positively oriented under a
jitter of at most 0.1 h per coordinate (the smallest Kuhn-tet altitude is h/sqrt(2)).

Node ordering follows ``Mesh::loadFromFile`` (``Mesh.cpp:803-878``): boundary (wall)
nodes first, then fluid nodes.  Wall nodes are un-jittered, flagged boundary, fixed
(``IC.BoundaryFixed = true``) and Dirichlet with v = 0 (``BoundaryV`` returns 0 in
``examples/2D/damBreakKoshizuka/damBreakKoshizukaIncomp.lua:83-85``).

Element-local node order within a cell is fixed, so two calls with the same
arguments return bit-identical arrays.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

# Node flag bits shared with include/pfem_gpu.h (PFEM_NODE_*).
BOUND = 1
FREE = 2  # derived from connectivity by the library (a node that belongs to no element)
ONFS = 4
FIXED = 8
DIRICHLET = 16


@dataclass
class PfemMesh:
    """A P1 mesh in the layouts the C-ABI takes.

    pos   : (dim, N) float64, field-blocked like the reference's q vectors
            (``StatesFromToQ.hpp:21-34``: entry (d, n) is q[n + d*N]).
    conn  : (E, dim+1) int32, element-local node order as ``Element::m_nodesIndexes``.
    flags : (N,) uint8, PFEM_NODE_* bits.
    surface : (N,) float64, height of the free surface above each node's column
            (used for hydrostatic initial conditions).
    """

    dim: int
    pos: np.ndarray
    conn: np.ndarray
    flags: np.ndarray
    h: float
    name: str = ""
    surface: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.pos.shape[1])

    @property
    def n_elems(self) -> int:
        return int(self.conn.shape[0])

    def free_mask(self) -> np.ndarray:
        used = np.zeros(self.n_nodes, dtype=bool)
        used[self.conn.reshape(-1)] = True
        return ~used


def _kuhn_tets() -> list[tuple[tuple[int, int, int], ...]]:
    """Six tets of the unit cube sharing the diagonal (0,0,0)-(1,1,1), positively oriented."""
    tets = []
    for perm in itertools.permutations(range(3)):
        a, b, _ = perm
        v0 = (0, 0, 0)
        v1 = [0, 0, 0]
        v1[a] = 1
        v2 = list(v1)
        v2[b] = 1
        v3 = (1, 1, 1)
        # det[e_a, e_a+e_b, e_a+e_b+e_c] = sign(perm)
        inv = sum(1 for i in range(3) for j in range(i + 1, 3) if perm[i] > perm[j])
        if inv % 2 == 0:
            tets.append((v0, tuple(v1), tuple(v2), v3))
        else:
            tets.append((v0, tuple(v2), tuple(v1), v3))
    return tets


_KUHN = _kuhn_tets()
_TRI = (((0, 0), (1, 0), (1, 1)), ((0, 0), (1, 1), (0, 1)))


def _jitter(rng: np.random.Generator, shape, h: float) -> np.ndarray:
    return rng.uniform(-0.1 * h, 0.1 * h, size=shape)


def _lattice_2d(nx: int, ny: int, h: float, is_wall, rng, jitter=True):
    """Nodes of an (nx+1) x (ny+1) lattice, walls first; two triangles per cell."""
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    ii = ii.reshape(-1)
    jj = jj.reshape(-1)
    wall = is_wall(ii, jj)
    order = np.concatenate([np.nonzero(wall)[0], np.nonzero(~wall)[0]])
    lat2node = np.empty(ii.size, dtype=np.int64)
    lat2node[order] = np.arange(ii.size)
    pos = np.stack([ii * h, jj * h]).astype(np.float64)[:, order]
    nwall = int(wall.sum())
    if jitter:
        pos[:, nwall:] += _jitter(rng, (2, ii.size - nwall), h)
    flags = np.zeros(ii.size, dtype=np.uint8)
    flags[:nwall] = BOUND | FIXED | DIRICHLET
    cj, ci = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ci = ci.reshape(-1)
    cj = cj.reshape(-1)
    elems = []
    for tri in _TRI:
        elems.append(np.stack([lat2node[(cj + dj) * (nx + 1) + (ci + di)] for (di, dj) in tri], axis=1))
    conn = np.stack(elems, axis=1).reshape(-1, 3).astype(np.int32)
    return pos, conn, flags, nwall, lat2node


def dambreak_2d(h=0.05, col_w=1.0, col_h=2.0, tank_w=4.0, tank_h=3.0, seed=0) -> PfemMesh:
    """2D dam break: a col_w x col_h fluid column in the corner of a tank_w x tank_h tank.

    Tank wall nodes sit every h along the bottom and both sides; those that touch
    no element are free nodes (identity rows in the PSPG system).  At the default
    sizes: 201 wall nodes + 800 fluid nodes = 1001 nodes, 1600 triangles.
    """
    rng = np.random.default_rng(seed)
    nx = int(round(col_w / h))
    ny = int(round(col_h / h))
    ntx = int(round(tank_w / h))
    nty = int(round(tank_h / h))
    # wall nodes: bottom (x = 0..tank_w), left (y = h..tank_h), right (y = h..tank_h)
    bottom = np.stack([np.arange(ntx + 1) * h, np.zeros(ntx + 1)])
    left = np.stack([np.zeros(nty), np.arange(1, nty + 1) * h])
    right = np.stack([np.full(nty, ntx * h), np.arange(1, nty + 1) * h])
    walls = np.concatenate([bottom, left, right], axis=1)
    nwall = walls.shape[1]
    fj, fi = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    fluid = np.stack([fi.reshape(-1) * h, fj.reshape(-1) * h]).astype(np.float64)
    fluid += _jitter(rng, fluid.shape, h)
    pos = np.concatenate([walls, fluid], axis=1)

    def lat(i, j):
        i = np.asarray(i)
        j = np.asarray(j)
        return np.where(j == 0, i, np.where(i == 0, ntx + 1 + (j - 1), nwall + (j - 1) * nx + (i - 1)))

    cj, ci = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ci = ci.reshape(-1)
    cj = cj.reshape(-1)
    elems = [np.stack([lat(ci + di, cj + dj) for (di, dj) in tri], axis=1) for tri in _TRI]
    conn = np.stack(elems, axis=1).reshape(-1, 3).astype(np.int32)
    flags = np.zeros(pos.shape[1], dtype=np.uint8)
    flags[:nwall] = BOUND | FIXED | DIRICHLET
    surface = np.full(pos.shape[1], col_h)
    m = PfemMesh(2, pos, conn, flags, h, name=f"dambreak2d_h{h}", surface=surface)
    m.meta.update(nx=nx, ny=ny, seed=seed)
    return m


def _lattice_3d(nx, ny, nz, h, is_wall, rng, cell_mask=None, jitter=True):
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    ii = ii.reshape(-1)
    jj = jj.reshape(-1)
    kk = kk.reshape(-1)
    nlat = ii.size
    # cells present
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ci = ci.reshape(-1)
    cj = cj.reshape(-1)
    ck = ck.reshape(-1)
    if cell_mask is not None:
        keep = cell_mask(ci, cj, ck)
        ci, cj, ck = ci[keep], cj[keep], ck[keep]

    def lid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    used = np.zeros(nlat, dtype=bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                used[lid(ci + di, cj + dj, ck + dk)] = True
    wall = is_wall(ii, jj, kk) & used
    fluid = (~wall) & used
    order = np.concatenate([np.nonzero(wall)[0], np.nonzero(fluid)[0]])
    lat2node = np.full(nlat, -1, dtype=np.int64)
    lat2node[order] = np.arange(order.size)
    pos = np.stack([ii * h, jj * h, kk * h]).astype(np.float64)[:, order]
    nwall = int(wall.sum())
    if jitter:
        pos[:, nwall:] += _jitter(rng, (3, order.size - nwall), h)
    flags = np.zeros(order.size, dtype=np.uint8)
    flags[:nwall] = BOUND | FIXED | DIRICHLET
    elems = []
    for tet in _KUHN:
        elems.append(np.stack([lat2node[lid(ci + a, cj + b, ck + c)] for (a, b, c) in tet], axis=1))
    conn = np.stack(elems, axis=1).reshape(-1, 4).astype(np.int32)
    return pos, conn, flags, order, (ii, jj, kk)


def dambreak_3d(h=0.01, col=(0.5, 0.5, 1.0), seed=0) -> PfemMesh:
    """3D dam break column of col[0] x col[1] x col[2] in the corner (x=0, y=0, z=0) of a box.

    Lattice nodes with i=0, j=0 or k=0 are wall nodes; the box's unwetted walls are
    omitted (they would be free nodes with identity rows).  At h = 0.01 this gives
    N = 51*51*101 = 262,701 nodes and E = 50*50*100*6 = 1,500,000 tets.
    """
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(round(c / h)) for c in col)
    pos, conn, flags, _, _ = _lattice_3d(nx, ny, nz, h, lambda i, j, k: (i == 0) | (j == 0) | (k == 0), rng)
    surface = np.full(pos.shape[1], nz * h)
    m = PfemMesh(3, pos, conn, flags, h, name=f"dambreak3d_{nx}x{ny}x{nz}", surface=surface)
    m.meta.update(nx=nx, ny=ny, nz=nz, seed=seed)
    return m


def randomwalk_3d(h=0.01, ncol=100, mean_cells=100, seed=0) -> PfemMesh:
    """3D column field: ncol x ncol cell columns whose heights follow a seeded +-1-cell random walk.

    Height of cell column (i, j): H(i, 0) walks along i from mean_cells, then H(i, j)
    walks along j from H(i, 0); heights are clipped to [mean/2, 3*mean/2].
    Walls at i=0, j=0, k=0.  At the defaults N ~ 1.0M and E ~ 6M.
    """
    rng = np.random.default_rng(seed)
    steps0 = rng.integers(-1, 2, size=ncol)
    steps0[0] = 0
    base = mean_cells + np.cumsum(steps0)
    steps = rng.integers(-1, 2, size=(ncol, ncol))
    steps[:, 0] = 0
    H = base[:, None] + np.cumsum(steps, axis=1)  # H[i, j]
    H = np.clip(H, mean_cells // 2, (3 * mean_cells) // 2)
    nz = int(H.max())

    def cell_mask(ci, cj, ck):
        return ck < H[ci, cj]

    pos, conn, flags, order, (ii, jj, kk) = _lattice_3d(ncol, ncol, nz, h, lambda i, j, k: (i == 0) | (j == 0) | (k == 0),
                                                         rng, cell_mask=cell_mask)
    # free-surface height over each node: max height of the (up to 4) cell columns around it
    Hp = np.zeros((ncol + 2, ncol + 2), dtype=np.int64)
    Hp[1:-1, 1:-1] = H
    li = ii[order]
    lj = jj[order]
    hn = np.maximum.reduce([Hp[li, lj], Hp[li + 1, lj], Hp[li, lj + 1], Hp[li + 1, lj + 1]])
    m = PfemMesh(3, pos, conn, flags, h, name=f"randomwalk3d_{ncol}", surface=hn * h)
    m.meta.update(ncol=ncol, nz=nz, seed=seed, mean_height=float(H.mean()))
    return m


def tank_3d(h=0.01, size=(2.0, 2.0, 1.0), fill=0.5, seed=0) -> PfemMesh:
    """3D sloshing tank of size[0] x size[1] filled to fill*size[2]; walls on the 4 sides and bottom.

    At h = 0.01: a 201 x 201 x 51 lattice, N = 2,060,451 nodes, E = 12,000,000 tets.
    """
    rng = np.random.default_rng(seed)
    nx = int(round(size[0] / h))
    ny = int(round(size[1] / h))
    nz = int(round(fill * size[2] / h))
    pos, conn, flags, _, _ = _lattice_3d(
        nx, ny, nz, h, lambda i, j, k: (i == 0) | (j == 0) | (k == 0) | (i == nx) | (j == ny), rng)
    surface = np.full(pos.shape[1], nz * h)
    m = PfemMesh(3, pos, conn, flags, h, name=f"tank3d_{nx}x{ny}x{nz}", surface=surface)
    m.meta.update(nx=nx, ny=ny, nz=nz, seed=seed)
    return m


def tank_2d(h=0.02, size=(2.0, 1.0), fill=0.5, seed=0) -> PfemMesh:
    """2D sloshing tank (walls left, right, bottom) filled to fill*size[1]."""
    rng = np.random.default_rng(seed)
    nx = int(round(size[0] / h))
    ny = int(round(fill * size[1] / h))
    pos, conn, flags, _, _ = _lattice_2d(nx, ny, h, lambda i, j: (i == 0) | (j == 0) | (i == nx), rng)
    surface = np.full(pos.shape[1], ny * h)
    m = PfemMesh(2, pos, conn, flags, h, name=f"tank2d_{nx}x{ny}", surface=surface)
    m.meta.update(nx=nx, ny=ny, seed=seed)
    return m


def column_2d_nodes(target_nodes: int, h=0.01, seed=0) -> PfemMesh:
    """2D dam-break column with aspect 1:2 sized to about target_nodes nodes (walls at x=0, y=0)."""
    nx = max(2, int(round((math.sqrt(1 + 8 * target_nodes) - 3) / 4)))
    ny = 2 * nx
    rng = np.random.default_rng(seed)
    pos, conn, flags, _, _ = _lattice_2d(nx, ny, h, lambda i, j: (i == 0) | (j == 0), rng)
    surface = np.full(pos.shape[1], ny * h)
    m = PfemMesh(2, pos, conn, flags, h, name=f"column2d_{nx}x{ny}", surface=surface)
    m.meta.update(nx=nx, ny=ny, seed=seed)
    return m


def column_3d_nodes(target_nodes: int, h=0.01, seed=0) -> PfemMesh:
    """3D dam-break column n x n x 2n sized to about target_nodes nodes (walls at i=0, j=0, k=0)."""
    n = max(2, int(round((target_nodes / 2.0) ** (1.0 / 3.0))) - 1)
    rng = np.random.default_rng(seed)
    pos, conn, flags, _, _ = _lattice_3d(n, n, 2 * n, h, lambda i, j, k: (i == 0) | (j == 0) | (k == 0), rng)
    surface = np.full(pos.shape[1], 2 * n * h)
    m = PfemMesh(3, pos, conn, flags, h, name=f"column3d_{n}", surface=surface)
    m.meta.update(n=n, seed=seed)
    return m


def tank_nodes(dim: int, target_nodes: int, h=0.01, seed=0) -> PfemMesh:
    """Sloshing tank (width 2x depth of fluid) of about target_nodes nodes."""
    if dim == 2:
        ny = max(2, int(round(math.sqrt(target_nodes / 2.0))) - 1)
        nx = 2 * ny
        rng = np.random.default_rng(seed)
        pos, conn, flags, _, _ = _lattice_2d(nx, ny, h, lambda i, j: (i == 0) | (j == 0) | (i == nx), rng)
        surface = np.full(pos.shape[1], ny * h)
        return PfemMesh(2, pos, conn, flags, h, name=f"tank2d_{nx}x{ny}", surface=surface, meta=dict(nx=nx, ny=ny))
    nz = max(2, int(round((target_nodes / 4.0) ** (1.0 / 3.0))) - 1)
    nx = 2 * nz
    rng = np.random.default_rng(seed)
    pos, conn, flags, _, _ = _lattice_3d(
        nx, nx, nz, h, lambda i, j, k: (i == 0) | (j == 0) | (k == 0) | (i == nx) | (j == nx), rng)
    surface = np.full(pos.shape[1], nz * h)
    return PfemMesh(3, pos, conn, flags, h, name=f"tank3d_{nx}x{nx}x{nz}", surface=surface, meta=dict(nx=nx, nz=nz))


def cavity(dim: int, n: int, h=None, seed=0) -> PfemMesh:
    """Closed square / cubic cavity of side 1 on an n^dim-cell lattice, every wall node boundary,
    fixed and Dirichlet (velocity): the lid-driven and heated-cavity cases of the heat,
    fractional-step and Boussinesq solvers.  meta["top"], meta["bottom"], meta["lid"] mark the
    top wall, the bottom wall, and the top nodes off the side walls."""
    h = 1.0 / n if h is None else h
    rng = np.random.default_rng(seed)
    if dim == 2:
        pos, conn, flags, _, _ = _lattice_2d(n, n, h, lambda i, j: (i == 0) | (j == 0) | (i == n) | (j == n), rng,
                                             jitter=True)
    else:
        pos, conn, flags, _, _ = _lattice_3d(
            n, n, n, h, lambda i, j, k: (i == 0) | (j == 0) | (k == 0) | (i == n) | (j == n) | (k == n), rng)
    m = PfemMesh(dim, pos, conn, flags, h, name=f"cavity{dim}d_{n}")
    top_z = n * h
    z = pos[dim - 1]
    side = np.zeros(pos.shape[1], dtype=bool)
    for k in range(dim - 1):
        side |= (np.abs(pos[k]) < 1e-12) | (np.abs(pos[k] - top_z) < 1e-12)
    top = np.abs(z - top_z) < 1e-12
    bottom = np.abs(z) < 1e-12
    m.meta.update(n=n, seed=seed, top=top, bottom=bottom, lid=top & ~side)
    return m


# ----------------------------------------------------------------------------------------
# Initial conditions (the examples' Problem.IC:initStates, evaluated in numpy)
# ----------------------------------------------------------------------------------------

def pspg_hydrostatic(mesh: PfemMesh, rho=1000.0, g=9.81) -> np.ndarray:
    """q = [u, v, (w), p] field-blocked; v = 0, p = rho*g*(H - z) below the surface, as
    ``damBreakKoshizukaIncomp.lua:71-81``."""
    d = mesh.dim
    N = mesh.n_nodes
    q = np.zeros((d + 1) * N)
    z = mesh.pos[d - 1]
    q[d * N:] = np.maximum(rho * g * (mesh.surface - z), 0.0)
    return q


def wc_tait_hydrostatic(mesh: PfemMesh, K0=2.2e9, K0p=7.0, rho_star=1000.0, g=9.81, u0=0.1):
    """WC state (v, p, rho, a) with the Tait-consistent hydrostatic column of
    ``examples/2D/sloshing/sloshingComp.lua:71-83`` and v_x = u0*sin(pi*x) on fluid nodes."""
    d = mesh.dim
    N = mesh.n_nodes
    z = mesh.pos[d - 1]
    depth = np.maximum(mesh.surface - z, 0.0)
    rho = rho_star * ((K0p - 1.0) / K0 * rho_star * g * depth + 1.0) ** (1.0 / (K0p - 1.0))
    p = K0 / K0p * ((rho / rho_star) ** K0p - 1.0)
    v = np.zeros((d, N))
    moving = (mesh.flags & BOUND) == 0
    v[0, moving] = u0 * np.sin(math.pi * mesh.pos[0, moving])
    a = np.zeros((d, N))
    return v, p, rho, a
