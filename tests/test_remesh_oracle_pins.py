"""Pins of the remeshing oracle (oracle/remesh_oracle.py) by what the geometry and the
reference's rules fix on hand-made inputs, each case also run through the library's mesher:

* Qhull's triangulation has empty circumcircles (brute force, exact arithmetic);
* a triangulated disc-like region has 2V - B - 2 triangles (Euler), its boundary nodes are the
  free surface and its boundary edges the facets;
* a gap wider than 2 alpha hchar is never bridged;
* simplices whose nodes are all boundary nodes are dropped when no interior simplex around them
  reaches a fluid node (a wall with no fluid meshes to nothing);
* addNodes puts one node at the centroid of an element larger than omega hchar^dim with the
  mean of its states; removeNodes drops the second of two close fluid nodes, never a wall node;
  checkBoundingBox drops nodes strictly outside; deleteFlyingNodes drops isolated fluid nodes,
  not isolated wall nodes.
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import pfem_amd
from oracle import remesh_oracle as RO

BOUND, FREE, ONFS = RO.BOUND, RO.FREE, RO.ONFS


def run_both(pos_nm, flags, prm: RO.Params, states=None):
    """Triangulate with the oracle and the library; return both (node-major)."""
    o = RO.triangulate(RO.from_cloud(pos_nm.T, flags, None if states is None else states.T), prm)
    m = pfem_amd.Mesher(prm.dim, hchar=prm.hchar, alpha=prm.alpha, omega=prm.omega, gamma=prm.gamma,
                        boundingBox=prm.bounding_box, deleteFlyingNodes=prm.delete_flying_nodes)
    m.set_nodes(pos_nm.T.copy(), flags, None if states is None else states.T.copy())
    m.triangulate()
    np.testing.assert_array_equal(m.elements(), o.elems)
    return o, m


def jittered_block(nx, ny, h, seed, x0=0.0):
    rng = np.random.default_rng(seed)
    P = np.array(list(itertools.product(range(nx), range(ny))), dtype=float) * h
    P += rng.uniform(-0.15 * h, 0.15 * h, size=P.shape)
    P[:, 0] += x0
    return P


def test_qhull_triangulation_has_empty_circumcircles():
    rng = np.random.default_rng(3)
    P = rng.random((40, 2))
    S, _ = RO._delaunay(P)
    F = [[Fraction(x) for x in p] for p in P]
    for s in S:
        a, b, c = (F[i] for i in s)
        for q in range(len(P)):
            if q in s:
                continue
            d = F[q]
            rows = [[p[0] - d[0], p[1] - d[1], (p[0] - d[0]) ** 2 + (p[1] - d[1]) ** 2] for p in (a, b, c)]
            det = (rows[0][0] * (rows[1][1] * rows[2][2] - rows[1][2] * rows[2][1])
                   - rows[0][1] * (rows[1][0] * rows[2][2] - rows[1][2] * rows[2][0])
                   + rows[0][2] * (rows[1][0] * rows[2][1] - rows[1][1] * rows[2][0]))
            assert det < 0   # d outside the circle of the counterclockwise triangle a, b, c


def test_block_counts_free_surface_and_facets():
    nx, ny, h = 9, 7, 0.1
    P = jittered_block(nx, ny, h, 1)
    flags = np.zeros(len(P), np.uint8)
    o, m = run_both(P, flags, RO.Params(2, h, alpha=1.2))
    V = len(P)
    border = np.array([i in (0, nx - 1) or j in (0, ny - 1) for i, j in itertools.product(range(nx), range(ny))])
    B = int(border.sum())
    assert len(o.elems) == 2 * V - B - 2
    np.testing.assert_array_equal((o.flags & ONFS) != 0, border)
    assert len(o.facets) == B
    # every facet is an edge of its element, the out node the element's third node
    for f, out, e in zip(o.facets, o.facet_out, o.facet_elem):
        assert set(f) | {out} == set(o.elems[e])
    _, _, fl, _ = m.nodes()
    np.testing.assert_array_equal(fl, o.flags)


def test_gap_is_not_bridged():
    h = 0.1
    A = jittered_block(6, 6, h, 2)
    B = jittered_block(6, 6, h, 3, x0=5 * h + 2.6 * 1.2 * h)   # gap > 2 alpha h between the blocks
    P = np.vstack([A, B])
    o, _ = run_both(P, np.zeros(len(P), np.uint8), RO.Params(2, h, alpha=1.2))
    side = np.arange(len(P)) >= len(A)
    for t in o.elems:
        assert side[t].all() or not side[t].any()
    assert len(o.elems) == 2 * (2 * 36 - 20 - 2)


def test_walls_without_fluid_mesh_to_nothing_next_to_fluid_they_stay():
    h = 0.1
    wall = jittered_block(8, 3, h, 4)
    flags = np.full(len(wall), BOUND, np.uint8)
    with pytest.raises(RuntimeError):
        RO.triangulate(RO.from_cloud(wall.T, flags), RO.Params(2, h, alpha=1.2))
    m = pfem_amd.Mesher(2, hchar=h, alpha=1.2)
    m.set_nodes(wall.T.copy(), flags)
    with pytest.raises(pfem_amd.PfemError):
        m.triangulate()
    # one fluid node above the wall's middle: only all-boundary triangles whose nodes have an
    # interior triangle reaching it survive, plus the triangles holding it
    fluid = np.array([[3.5 * h, 2.8 * h]])
    P = np.vstack([wall, fluid])
    fl = np.concatenate([flags, [0]]).astype(np.uint8)
    o, _ = run_both(P, fl, RO.Params(2, h, alpha=1.2))
    fi = len(wall)
    S, _ = RO._delaunay(P)
    alpha2 = (1.2 * h) ** 2
    interior = [RO.circumradius2(P, s) <= alpha2 for s in S]
    reach = set()  # nodes of interior triangles holding the fluid node
    for s, ok in zip(S, interior):
        if ok and fi in s:
            reach.update(int(v) for v in s)
    for t in o.elems:
        assert fi in t or any(v in reach for v in t)
    assert any(fi not in t for t in o.elems)


def test_add_nodes_centroid_and_states():
    h = 0.1
    P = np.array([[0, 0], [1, 0], [0, 1]], dtype=float) * 0.5
    st = np.array([[1.0, 2.0], [3.0, 5.0], [8.0, 13.0]])
    prm = RO.Params(2, h, alpha=10.0, omega=0.5)     # element area 0.125 > 0.5 h^2
    o = RO.triangulate(RO.from_cloud(P.T, np.zeros(3, np.uint8), st.T), prm)
    RO.add_nodes(o, prm)
    assert len(o.pos) == 4
    np.testing.assert_array_equal(o.pos[3], [(0 + 0.5 + 0) / 3, (0 + 0 + 0.5) / 3])
    np.testing.assert_array_equal(o.states[3], [(1 + 3 + 8) / 3, (2 + 5 + 13) / 3])
    assert o.tags[3] == -1 and o.flags[3] == FREE
    m = pfem_amd.Mesher(2, hchar=h, alpha=10.0, omega=0.5)
    m.set_nodes(P.T.copy(), np.zeros(3, np.uint8), st.T.copy())
    m.triangulate()
    m.remesh()
    pos, s2, _, tg = m.nodes()
    np.testing.assert_array_equal(pos.T[3], o.pos[3])
    np.testing.assert_array_equal(s2.T[3], o.states[3])
    assert tg[3] == -1 and len(m.elements()) == 3


def test_remove_nodes_and_bounding_box_and_flying():
    h = 0.1
    P = jittered_block(6, 6, h, 5)
    P = np.vstack([P, P[14] + [0.02 * h, 0.01 * h]])          # a fluid node right next to node 14
    flags = np.zeros(len(P), np.uint8)
    flags[0] = BOUND
    P = np.vstack([P, P[0] + [0.01 * h, 0.0]])                # a fluid node next to the wall node 0
    flags = np.concatenate([flags, [0]]).astype(np.uint8)
    prm = RO.Params(2, h, alpha=1.2, gamma=0.3)
    o = RO.triangulate(RO.from_cloud(P.T, flags), prm)
    r = RO.remove_nodes(o, prm)
    # 14 comes first in node order: it is touched, its close neighbour (36) goes; node 37 is
    # close to the wall node 0, which is never deleted: 37 goes
    assert len(r.pos) == len(P) - 2
    np.testing.assert_array_equal(r.pos, np.delete(P, [36, 37], axis=0))
    prm_bb = RO.Params(2, h, alpha=1.2, bounding_box=(-1.0, -1.0, 0.3, 1.0))
    b = RO.check_bounding_box(o, prm_bb)
    np.testing.assert_array_equal(b.pos, P[P[:, 0] <= 0.3])
    # an isolated fluid node is flying: deleted; an isolated wall node stays (free)
    Q = np.vstack([jittered_block(5, 5, h, 6), [[3.0, 3.0], [-3.0, 3.0]]])
    fq = np.zeros(len(Q), np.uint8)
    fq[-1] = BOUND
    o2, m2 = run_both(Q, fq, RO.Params(2, h, alpha=1.2, delete_flying_nodes=True))
    assert len(o2.pos) == len(Q) - 1
    np.testing.assert_array_equal(o2.pos[-1], [-3.0, 3.0])
    assert o2.flags[-1] & FREE
    # the library remeshes the same way
    m = pfem_amd.Mesher(2, hchar=h, alpha=1.2, gamma=0.3)
    m.set_nodes(P.T.copy(), flags)
    m.triangulate()
    m.remesh()
    pos, _, _, _ = m.nodes()
    np.testing.assert_array_equal(pos.T, np.delete(P, [36, 37], axis=0))
