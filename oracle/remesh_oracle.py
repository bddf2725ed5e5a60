"""CPU oracle of the reference's remeshing (TEST INFRASTRUCTURE ONLY: used by tests/, never by the
library, the smoke test's product path or bench.py's measured work).

A restatement, in numpy over scipy's Qhull Delaunay triangulation, of
  * Mesh::triangulateAlphaShape2D / 3D (srcs/mesh/Mesh2D.cpp:32-251, Mesh3D.cpp:36-286): the
    alpha-shape classification, the deletion of all-boundary simplices, the free-surface facets
    and flags, the node neighbour lists, and deleteFlyingNodes (Mesh.cpp:251-331);
  * Mesh::remesh (Mesh.cpp:919-926): laplacianSmoothingBoundaries (Mesh.cpp:580-760), addNodes
    (Mesh.cpp:27-147), removeNodes (Mesh.cpp:928-1017), checkBoundingBox (Mesh.cpp:149-207).

Qhull's triangulation is the Delaunay triangulation only for points in general position (no
four co-circular / five co-spherical points); the oracle is used on jittered clouds.  Element
order and node rotation follow the library's canonical form (sorted node tuples, smallest node
first with the orientation kept) so that added nodes are numbered, and their centroids summed,
in the same order; the reference's own order is CGAL's iteration order.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import Delaunay

BOUND, FREE, ONFS, FIXED, DIRICHLET, DIRICHLET_T = 1, 2, 4, 8, 16, 32

# faces opposite each vertex, oriented to see that vertex on their positive side
FACE = {2: ((1, 2), (2, 0), (0, 1)), 3: ((1, 3, 2), (0, 2, 3), (0, 3, 1), (0, 1, 2))}
_EVEN4 = [p for p in itertools.permutations(range(4))
          if sum(1 for i in range(4) for j in range(i + 1, 4) if p[i] > p[j]) % 2 == 0]


def canonical(t):
    t = list(t)
    if len(t) == 3:
        return min(tuple(t[r:] + t[:r]) for r in range(3))
    return min(tuple(t[i] for i in p) for p in _EVEN4)


def det_j(P, v):
    """Element::computeJ + computeDetJ (Element.cpp:15-86), P node-major."""
    if len(v) == 3:
        a, b, c = P[v[0]], P[v[1]], P[v[2]]
        J00, J01, J10, J11 = b[0] - a[0], c[0] - a[0], b[1] - a[1], c[1] - a[1]
        return J00 * J11 - J10 * J01
    a, b, c, d = (P[i] for i in v)
    J = [[b[k] - a[k], c[k] - a[k], d[k] - a[k]] for k in range(3)]
    return (J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1]
            - J[2][0] * J[1][1] * J[0][2] - J[2][1] * J[1][2] * J[0][0] - J[2][2] * J[1][0] * J[0][1])


def circumradius2(P, v):
    """Squared circumradius (the alpha value of a Delaunay simplex)."""
    A = P[list(v)]
    M = 2.0 * (A[1:] - A[0])
    rhs = (A[1:] ** 2).sum(1) - (A[0] ** 2).sum()
    c = np.linalg.solve(M, rhs)
    return float(((c - A[0]) ** 2).sum())


@dataclass
class MeshState:
    """The node list and triangulation of the reference's Mesh (node-major arrays)."""
    dim: int
    pos: np.ndarray            # N x dim
    states: np.ndarray         # N x nstates
    flags: np.ndarray          # N uint8 (BOUND, FIXED, DIRICHLET, DIRICHLET_T kept; FREE, ONFS computed)
    tags: np.ndarray           # N int32
    elems: np.ndarray = None   # E x (dim+1)
    facets: np.ndarray = None  # F x dim
    facet_out: np.ndarray = None
    facet_elem: np.ndarray = None
    nbrs: list = field(default_factory=list)   # node -> sorted neighbour list
    nelem: np.ndarray = None


@dataclass
class Params:
    dim: int
    hchar: float
    alpha: float = 1.0
    omega: float = 1e16
    gamma: float = 0.0
    add_on_fs: bool = True
    delete_flying_nodes: bool = False
    laplacian_smoothing_boundaries: bool = False
    bounding_box: tuple | None = None
    exclusion_zones: tuple = ()


def _delaunay(P):
    """Positively oriented Delaunay simplices and their neighbours (-1 across the hull)."""
    tri = Delaunay(P)
    S = tri.simplices.copy()
    Nb = tri.neighbors.copy()
    for s in range(len(S)):
        if det_j(P, S[s]) < 0:
            S[s][[0, 1]] = S[s][[1, 0]]
            Nb[s][[0, 1]] = Nb[s][[1, 0]]
    return S, Nb


def triangulate(m: MeshState, prm: Params) -> MeshState:
    d, nv = m.dim, m.dim + 1
    P = m.pos
    n = len(P)
    bound = (m.flags & BOUND) != 0
    S, Nb = _delaunay(P)
    alpha2 = prm.alpha * prm.alpha * prm.hchar * prm.hchar
    interior = np.array([circumradius2(P, s) <= alpha2 for s in S])
    incident = [[] for _ in range(n)]  # vertex -> incident finite simplices
    for si, s in enumerate(S):
        for v in s:
            incident[v].append(si)

    def check_deletion(si):
        """checkFaceDeletion (Mesh2D.cpp:58-118) / checkCellDeletion (Mesh3D.cpp:63-128)."""
        v = S[si]
        mean = [sum(P[j][k] for j in v) / nv for k in range(d)]
        for z in prm.exclusion_zones:
            if all(mean[k] > z[k] for k in range(d)) and all(mean[k] < z[d + k] for k in range(d)):
                return True
        if all(bound[j] for j in v):
            if d == 3 and det_j(P, v) / 6.0 < 0.02 * prm.hchar ** 3:
                return True
            for i in range(nv):
                neighbour_vh = set()
                for t in incident[v[i]]:
                    if interior[t]:
                        neighbour_vh.update(int(w) for w in S[t])
                for w in neighbour_vh:
                    if w in v:
                        continue
                    if not bound[w]:
                        return False
            return True
        return False

    keep = np.array([interior[si] and not check_deletion(si) for si in range(len(S))])
    if not keep.any():
        raise RuntimeError("Something went wrong while remeshing.")
    kept = sorted((canonical(S[si]), si) for si in range(len(S)) if keep[si])
    elem_of = {si: e for e, (_, si) in enumerate(kept)}
    elems = np.array([t for t, _ in kept], dtype=np.int32)
    flags = m.flags & ~np.uint8(ONFS | FREE)
    # free-surface facets: REGULAR faces (one side interior) whose interior side is kept
    fl = []
    for si in range(len(S)):
        if not keep[si]:
            continue
        for i in range(nv):
            t = Nb[si][i]
            if t >= 0 and interior[t]:
                continue
            f = [int(S[si][k]) for k in FACE[d][i]]
            if d == 3:
                r = int(np.argmin(f))
                f = f[r:] + f[:r]
            fl.append((tuple(f), int(S[si][i]), elem_of[si]))
    fl.sort()
    for f, _, _ in fl:
        if not all(bound[j] for j in f):
            for j in f:
                flags[j] |= ONFS
    nbrs = [set() for _ in range(n)]
    nelem = np.zeros(n, dtype=np.int64)
    for t in elems:
        for a in t:
            nelem[a] += 1
            nbrs[a].update(int(b) for b in t if b != a)
    flags[nelem == 0] |= FREE
    out = MeshState(d, P.copy(), m.states.copy(), flags, m.tags.copy(), elems,
                    np.array([f for f, _, _ in fl], dtype=np.int32).reshape(-1, d),
                    np.array([o for _, o, _ in fl], dtype=np.int32), np.array([e for _, _, e in fl], dtype=np.int32),
                    [sorted(s) for s in nbrs], nelem)
    if prm.delete_flying_nodes:
        out = delete_flying_nodes(out)
    return out


def _erase(m: MeshState, delete: np.ndarray, renumber: bool) -> MeshState:
    keep = ~delete
    newid = -np.ones(len(delete), dtype=np.int64)
    newid[keep] = np.arange(keep.sum())
    out = MeshState(m.dim, m.pos[keep].copy(), m.states[keep].copy(), m.flags[keep].copy(), m.tags[keep].copy())
    out.nelem = m.nelem[keep].copy() if m.nelem is not None else None
    if renumber:
        out.elems = newid[m.elems].astype(np.int32)
        out.facets = newid[m.facets].astype(np.int32)
        out.facet_out = newid[m.facet_out].astype(np.int32)
        out.facet_elem = m.facet_elem.copy()
        out.nbrs = [[int(newid[j]) for j in m.nbrs[i] if newid[j] >= 0] for i in np.nonzero(keep)[0]]
    return out


def delete_flying_nodes(m: MeshState) -> MeshState:
    """Mesh::deleteFlyingNodes (Mesh.cpp:251-331); facets are renumbered too (the reference leaves
    them stale)."""
    delete = (m.nelem == 0) & ((m.flags & BOUND) == 0)
    return _erase(m, delete, True) if delete.any() else m


def laplacian_smoothing_boundaries(m: MeshState, prm: Params):
    """Mesh::laplacianSmoothingBoundaries (Mesh.cpp:580-760), positions updated in place in
    element order."""
    if not prm.laplacian_smoothing_boundaries:
        return
    d, h, P = m.dim, prm.hchar, m.pos
    bound = (m.flags & BOUND) != 0
    for el in m.elems:
        if not any(bound[v] for v in el):
            continue
        facet = [int(v) for v in el if bound[v]]
        outs = [int(v) for v in el if not bound[v]]
        if len(facet) != d:
            continue
        o = outs[-1]
        nrm = np.zeros(3)
        if d == 2:
            (x0, y0), (x1, y1) = P[facet[0]], P[facet[1]]
            if x0 == x1:
                nrm[0] = 1
                dist = abs(x0 - P[o][0])
            elif y0 == y1:
                nrm[1] = 1
                dist = abs(y0 - P[o][1])
            else:
                if x0 * y1 - x1 * y0 == 0:   # a line through the origin: no a x + b y + 1 = 0 form
                    continue
                coeff = np.linalg.solve(np.array([[x0, y0], [x1, y1]]), np.array([-1.0, -1.0]))
                nrm[:2] = coeff
                dist = abs(coeff[0] * P[o][0] + coeff[1] * P[o][1] + 1) / np.hypot(coeff[0], coeff[1])
        else:
            A = P[facet]
            if A[0][0] == A[1][0] == A[2][0]:
                nrm[0] = 1
                dist = abs(A[0][0] - P[o][0])
            elif A[0][1] == A[1][1] == A[2][1]:
                nrm[1] = 1
                dist = abs(A[0][1] - P[o][1])
            elif A[0][2] == A[1][2] == A[2][2]:   # Mesh.cpp:662 compares A[1][1] with A[2][2] (corrected)
                nrm[2] = 1
                dist = abs(A[0][2] - P[o][2])
            else:
                if np.linalg.det(A) == 0:   # a plane through the origin: element left alone
                    continue
                coeff = np.linalg.solve(A, -np.ones(3))
                nrm[:] = coeff
                dist = abs(coeff @ P[o] + 1) / np.linalg.norm(coeff)
        if dist > h / 3:
            continue
        if m.flags[o] & ONFS:
            toward = P[o] - P[facet[0]]
            if d == 2:
                if nrm[0] * toward[0] + nrm[1] * toward[1] < 0:
                    nrm = -nrm
                nrm = nrm / np.hypot(nrm[0], nrm[1])
            else:
                nrm = nrm / np.linalg.norm(nrm)
                if toward @ nrm < 0:
                    nrm = -nrm
            for k in range(d):
                P[o][k] += h / 3 * nrm[k]
        else:
            nb = m.nbrs[o]
            acc = np.zeros(d)
            for j in nb:
                for k in range(d):
                    acc[k] += P[j][k]
            P[o] = acc / float(len(nb))


def add_nodes(m: MeshState, prm: Params) -> int:
    """Mesh::addNodes (Mesh.cpp:27-147)."""
    d, nv = m.dim, m.dim + 1
    limit = prm.omega * prm.hchar ** d
    ref = 0.5 if d == 2 else 1.0 / 6.0
    new_pos, new_st = [], []
    onfs = (m.flags & ONFS) != 0
    for el in m.elems:
        size = det_j(m.pos, el) * ref
        if size > limit and (prm.add_on_fs or not any(onfs[v] for v in el)):
            p = []
            for k in range(d):
                s = 0.0
                for v in el:
                    s += m.pos[v][k]
                p.append(s / nv)
            st = []
            for q in range(m.states.shape[1]):
                s = 0.0
                for v in el:
                    s += m.states[v][q]
                st.append(s / nv)
            new_pos.append(p)
            new_st.append(st)
    if new_pos:
        k = len(new_pos)
        m.pos = np.vstack([m.pos, np.array(new_pos)])
        m.states = np.vstack([m.states, np.array(new_st).reshape(k, m.states.shape[1])])
        m.flags = np.concatenate([m.flags, np.full(k, FREE, dtype=np.uint8)])
        m.tags = np.concatenate([m.tags, np.full(k, -1, dtype=np.int32)])
        m.nelem = np.concatenate([m.nelem, np.zeros(k, dtype=m.nelem.dtype)])
        m.nbrs = list(m.nbrs) + [[] for _ in range(k)]
    return len(new_pos)


def remove_nodes(m: MeshState, prm: Params) -> MeshState:
    """Mesh::removeNodes (Mesh.cpp:928-1017)."""
    n, d = len(m.pos), m.dim
    touched = np.zeros(n, dtype=bool)
    delete = np.zeros(n, dtype=bool)
    limit = prm.gamma * prm.hchar
    onfs = (m.flags & ONFS) != 0
    bound = (m.flags & BOUND) != 0
    for i in range(n):
        if delete[i] or m.nelem[i] == 0:
            continue
        for j in m.nbrs[i]:
            dd = 0.0
            for k in range(d):
                dd += (m.pos[i][k] - m.pos[j][k]) * (m.pos[i][k] - m.pos[j][k])
            dd = np.sqrt(dd)
            fact = 0.5 if (onfs[i] and onfs[j]) else 1.0
            if dd <= fact * limit:
                if touched[j]:
                    if bound[i] or (onfs[i] and not onfs[j]):
                        continue
                    delete[i] = True
                else:
                    if bound[j] or (onfs[j] and not onfs[i]):
                        continue
                    touched[i] = True
                    delete[j] = True
    return _erase(m, delete, False) if delete.any() else m


def check_bounding_box(m: MeshState, prm: Params) -> MeshState:
    """Mesh::checkBoundingBox (Mesh.cpp:149-207)."""
    if prm.bounding_box is None:
        return m
    d = m.dim
    bb = np.asarray(prm.bounding_box)
    out = ((m.pos < bb[:d]) | (m.pos > bb[d:])).any(axis=1)
    return _erase(m, out, False) if out.any() else m


def remesh(m: MeshState, prm: Params) -> MeshState:
    """Mesh::remesh (Mesh.cpp:919-926)."""
    m = MeshState(m.dim, m.pos.copy(), m.states.copy(), m.flags.copy(), m.tags.copy(), m.elems, m.facets,
                  m.facet_out, m.facet_elem, [list(x) for x in m.nbrs], m.nelem.copy())
    laplacian_smoothing_boundaries(m, prm)
    add_nodes(m, prm)
    m = remove_nodes(m, prm)
    m = check_bounding_box(m, prm)
    return triangulate(m, prm)


def from_cloud(pos_fb, flags, states_fb=None, tags=None) -> MeshState:
    """Node-major MeshState from field-blocked arrays (the C-ABI layout)."""
    pos = np.ascontiguousarray(np.asarray(pos_fb, dtype=np.float64).T)
    n = len(pos)
    st = np.zeros((n, 0)) if states_fb is None else np.ascontiguousarray(np.asarray(states_fb).reshape(-1, n).T)
    fl = np.asarray(flags, dtype=np.uint8) & np.uint8(BOUND | FIXED | DIRICHLET | DIRICHLET_T)
    tg = -np.ones(n, dtype=np.int32) if tags is None else np.asarray(tags, dtype=np.int32).copy()
    return MeshState(pos.shape[1], pos, st, fl, tg)
