"""Host remesher (pfem_mesher_*, mesher.cpp) against the remeshing oracle
(oracle/remesh_oracle.py) and against what the geometry fixes.  CPU only: the mesher does no
GPU work, so these run through the C-ABI library in any environment.

* The Delaunay triangulation equals Qhull's on points in general position (2D, 3D).
* On lattices, where every square / cube is co-circular / co-spherical, it is the perturbed
  Delaunay triangulation: no point lies inside any simplex's circumsphere under the lifting
  perturbation (each point's lift raised by eps^rank, the lexicographically largest point the
  most), evaluated here with exact rational cofactors; the simplices tile the hull.
* triangulate (alpha shape, all-boundary deletions, exclusion zones, free-surface facets and
  flags, flying nodes) and remesh (smoothing, addNodes, removeNodes, bounding box) equal the
  oracle's on jittered clouds with walls, element by element and bit for bit in the node data.
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import pfem_amd
from oracle import remesh_oracle as RO

BOUND, FREE, ONFS = RO.BOUND, RO.FREE, RO.ONFS


def tank_cloud(dim, h, size, fill, seed, jitter=0.15):
    """Nodes of an open-top tank: wall nodes on a lattice of spacing h, each moved within its
    wall(s) by a jitter that depends only on its lattice index (nodes on an edge move along the
    edge, corners stay), and a block of fluid nodes jittered in every direction.  No four nodes
    are co-circular, so Qhull's triangulation is the Delaunay one."""
    rng = np.random.default_rng(seed)
    L = [int(round(s / h)) for s in size]

    def wall_jitter(idx):
        r = np.random.default_rng(hash((seed,) + tuple(int(i) for i in idx)) & 0xFFFFFFFF)
        return r.uniform(-jitter * h, jitter * h, size=dim)

    walls = []
    for idx in itertools.product(*[range(k + 1) for k in L]):
        pinned = [idx[a] in (0, L[a]) for a in range(dim)]
        on_wall = any(pinned[:-1]) or idx[-1] == 0
        if not on_wall or (idx[-1] == L[-1] and not any(pinned[:-1])):
            continue
        j = wall_jitter(idx)
        walls.append([idx[a] * h + (0.0 if pinned[a] else j[a]) for a in range(dim)])
    walls = np.array(walls)
    nf = [int(round(f / h)) for f in fill]
    grid = np.array(list(itertools.product(*[range(1, k) for k in nf])), dtype=float) * h
    grid += rng.uniform(-jitter * h, jitter * h, size=grid.shape)
    pos = np.vstack([walls, grid])
    flags = np.concatenate([np.full(len(walls), BOUND, np.uint8), np.zeros(len(grid), np.uint8)])
    return pos.T.copy(), flags


def make_mesher(dim, **kw):
    return pfem_amd.Mesher(dim, **kw)


def lib_state(m):
    pos, st, fl, tg = m.nodes()
    conn = m.elements()
    fn, fo, fe = m.facets()
    return dict(pos=pos.T, states=st.T, flags=fl, tags=tg, elems=conn, facets=fn, facet_out=fo, facet_elem=fe)


def assert_same(lib, orc, exact=True):
    if exact:
        np.testing.assert_array_equal(lib["pos"], orc.pos)
        np.testing.assert_array_equal(lib["states"], orc.states)
    else:
        np.testing.assert_allclose(lib["pos"], orc.pos, rtol=0, atol=1e-12)
        np.testing.assert_allclose(lib["states"], orc.states, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(lib["flags"], orc.flags)
    np.testing.assert_array_equal(lib["tags"], orc.tags)
    np.testing.assert_array_equal(lib["elems"], orc.elems)
    np.testing.assert_array_equal(lib["facets"], orc.facets)
    np.testing.assert_array_equal(lib["facet_out"], orc.facet_out)
    np.testing.assert_array_equal(lib["facet_elem"], orc.facet_elem)


# ---- Delaunay --------------------------------------------------------------------------------
@pytest.mark.parametrize("dim,n", [(2, 40), (2, 3000), (3, 30), (3, 2500)])
def test_delaunay_equals_qhull_in_general_position(dim, n):
    rng = np.random.default_rng(dim * 1000 + n)
    P = rng.random((dim, n))
    m = make_mesher(dim, hchar=1.0, alpha=1e9)
    m.set_nodes(P, np.zeros(n, np.uint8))
    m.triangulate()
    E = m.elements()
    S, _ = RO._delaunay(P.T.copy())
    assert {tuple(t) for t in E} == {RO.canonical(s) for s in S}
    # positively oriented, canonical (smallest node first), sorted
    assert all(RO.det_j(P.T, t) > 0 for t in E)
    assert all(t[0] == t.min() for t in E)
    assert [tuple(t) for t in E] == sorted(tuple(t) for t in E)


def _det(M):
    """Exact determinant (fraction-free elimination over Fractions)."""
    M = [[Fraction(x) for x in r] for r in M]
    n, sign, prev = len(M), 1, Fraction(1)
    for k in range(n - 1):
        if M[k][k] == 0:
            for i in range(k + 1, n):
                if M[i][k] != 0:
                    M[k], M[i] = M[i], M[k]
                    sign = -sign
                    break
            else:
                return Fraction(0)
        for i in range(k + 1, n):
            for j in range(k + 1, n):
                M[i][j] = (M[i][j] * M[k][k] - M[i][k] * M[k][j]) / prev
        prev = M[k][k]
    return sign * M[-1][-1]


def _perturbed_inside(P, simplex, q):
    """q strictly inside the circumsphere of `simplex` with every lift |p|^2 raised by eps^rank
    (rank 1 for the lexicographically largest point): the lifted determinant is linear in the
    lift column, so its sign is that of the unperturbed value, else of the first nonzero cofactor
    of a lift entry in rank order.  The sign meaning 'inside' is read off the simplex's centroid."""
    D = len(P[0])
    ids = list(simplex) + [q]
    rows = [[*P[i], sum(Fraction(c) ** 2 for c in P[i]), 1] for i in ids]
    cen = [sum(Fraction(P[i][k]) for i in simplex) / len(simplex) for k in range(D)]
    s_in = _det(rows[:-1] + [[*cen, sum(c * c for c in cen), 1]]) > 0
    d0 = _det(rows)
    if d0 != 0:
        return (d0 > 0) == s_in
    for r in sorted(range(len(ids)), key=lambda r: tuple(P[ids[r]]), reverse=True):
        c = (-1) ** (r + D) * _det([row[:D] + row[D + 1:] for j, row in enumerate(rows) if j != r])
        if c != 0:
            return (c > 0) == s_in
    raise AssertionError("distinct points always have a nonzero coefficient")


def _check_perturbed_delaunay(P, E):
    pts = [tuple(float(x) for x in p) for p in P]
    for s in E:
        for q in range(len(pts)):
            if q not in s:
                assert not _perturbed_inside(pts, s, q), (s, q)
    # a tiling of the convex hull: every face shared by two simplices or on the hull, volumes add up
    from scipy.spatial import ConvexHull
    d = len(pts[0])
    vol = sum(abs(_det([[Fraction(P[v][k]) - Fraction(P[s[0]][k]) for k in range(d)] for v in s[1:]])) for s in E)
    vol /= (2 if d == 2 else 6)
    assert float(vol) == pytest.approx(ConvexHull(P).volume, rel=1e-12)
    faces = {}
    for s in E:
        for i in range(d + 1):
            f = tuple(sorted(np.delete(s, i)))
            faces[f] = faces.get(f, 0) + 1
    assert max(faces.values()) <= 2


@pytest.mark.parametrize("shape,extra", [((7, 5), 0), ((6, 6), 6), ((4, 4, 3), 0), ((3, 3, 3), 5)])
def test_lattice_triangulation_is_the_perturbed_delaunay(shape, extra):
    """Every unit square / cube is co-circular / co-spherical; extra points on a 1/4 sub-lattice
    add more ties (and collinear / coplanar hull points)."""
    dim = len(shape)
    P = np.array(list(itertools.product(*[range(k) for k in shape])), dtype=float)
    if extra:
        rng = np.random.default_rng(len(shape) * 10 + extra)
        sub = rng.integers(1, 4 * (min(shape) - 1), size=(extra, dim)) / 4.0
        P = np.unique(np.vstack([P, sub]), axis=0)
    m = make_mesher(dim, hchar=1.0, alpha=1e9)
    m.set_nodes(P.T.copy(), np.zeros(len(P), np.uint8))
    m.triangulate()
    E = m.elements()
    assert all(RO.det_j(P, t) > 0 for t in E)
    _check_perturbed_delaunay(P, [list(t) for t in E])


def test_ulp_perturbed_lattice_uses_exact_predicates():
    """A lattice moved by a few ulps: the floating-point filters cannot decide these signs."""
    rng = np.random.default_rng(5)
    P = np.array(list(itertools.product(range(5), range(5))), dtype=float) * 0.1
    P = P + rng.integers(-2, 3, size=P.shape) * np.spacing(np.maximum(np.abs(P), 0.1))
    m = make_mesher(2, hchar=1.0, alpha=1e9)
    m.set_nodes(P.T.copy(), np.zeros(len(P), np.uint8))
    m.triangulate()
    _check_perturbed_delaunay(P, [list(t) for t in m.elements()])


def test_duplicate_positions_are_not_inserted():
    P = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [1, 0], [0.4, 0.3]], dtype=float)
    m = make_mesher(2, hchar=1.0, alpha=1e9)
    m.set_nodes(P.T.copy(), np.zeros(len(P), np.uint8))
    m.triangulate()
    assert m.stats()["duplicates"] == 1
    _, _, fl, _ = m.nodes()
    assert fl[4] & FREE and not (fl[1] & FREE)


# ---- triangulate / remesh against the oracle ---------------------------------------------------
CASES = {
    "2d": dict(dim=2, h=0.05, size=(1.0, 0.8), fill=(0.5, 0.5)),
    "3d": dict(dim=3, h=0.1, size=(0.6, 0.4, 0.6), fill=(0.3, 0.4, 0.4)),
}


def _setup(name, seed=0, nstates=None, **prm):
    c = CASES[name]
    pos, flags = tank_cloud(c["dim"], c["h"], c["size"], c["fill"], seed)
    n = pos.shape[1]
    ns = (c["dim"] + 1) if nstates is None else nstates
    st = np.random.default_rng(seed + 1).normal(size=(ns, n))
    tags = np.where(flags & BOUND, 0, 1).astype(np.int32)
    kw = dict(hchar=c["h"], alpha=1.2)
    kw.update(prm)
    m = make_mesher(c["dim"], **kw)
    m.set_nodes(pos, flags, st, tags)
    o = RO.from_cloud(pos, flags, st, tags)
    p = RO.Params(c["dim"], c["h"], kw["alpha"], kw.get("omega", 1e16), kw.get("gamma", 0.0), kw.get("addOnFS", True),
                  kw.get("deleteFlyingNodes", False), kw.get("laplacianSmoothingBoundaries", False),
                  kw.get("boundingBox"), tuple(kw.get("exclusionZones", ())))
    return m, o, p


@pytest.mark.parametrize("name", ["2d", "3d"])
@pytest.mark.parametrize("flying", [False, True])
def test_triangulate_equals_oracle(name, flying):
    m, o, p = _setup(name, deleteFlyingNodes=flying)
    m.triangulate()
    ref = RO.triangulate(o, p)
    lib = lib_state(m)
    assert_same(lib, ref)
    # the free surface: the open top of the fluid block (its nodes and the walls they touch)
    fl = lib["flags"]
    assert (fl & ONFS).any() and not (fl & ONFS)[(fl & BOUND) == 0].all()
    # walls above the fluid hold no element: free unless deleted as flying (they are boundary)
    assert ((fl & FREE) != 0).sum() > 0
    if name == "3d":
        P = lib["pos"]
        for f, out in zip(lib["facets"], lib["facet_out"]):
            a, b, c = P[f]
            assert np.dot(np.cross(b - a, c - a), P[out] - a) > 0


def test_exclusion_zone_drops_elements():
    m, o, p = _setup("2d", exclusionZones=[(0.1, 0.1, 0.3, 0.3)])
    m.triangulate()
    assert_same(lib_state(m), RO.triangulate(o, p))
    pos = lib_state(m)["pos"]
    cen = pos[m.elements()].mean(axis=1)
    assert not ((cen > 0.1) & (cen < 0.3)).all(axis=1).any()


@pytest.mark.parametrize("name", ["2d", "3d"])
def test_remesh_cycles_equal_oracle(name):
    """Three remesh cycles on a cloud whose nodes move between them: addNodes (omega small enough
    to split the largest elements), removeNodes (gamma), checkBoundingBox, deleteFlyingNodes."""
    c = CASES[name]
    d = c["dim"]
    bb = [-1e-9] * d + [s + 1e-9 for s in c["size"]]
    bb[2 * d - 1] = c["fill"][-1] * 1.05   # the box top cuts the highest fluid nodes
    m, o, p = _setup(name, omega=0.55 if d == 2 else 0.14, gamma=0.5, boundingBox=bb, deleteFlyingNodes=True)
    m.triangulate()
    ref = RO.triangulate(o, p)
    rng = np.random.default_rng(7)
    added = removed = 0
    for cycle in range(3):
        lib = lib_state(m)
        n = len(lib["pos"])
        move = rng.normal(scale=0.12 * c["h"], size=(n, d))
        move[(lib["flags"] & BOUND) != 0] = 0
        pos = lib["pos"] + move
        st = lib["states"] + rng.normal(size=lib["states"].shape)
        m.set_positions_states(pos.T.copy(), st.T.copy())
        ref.pos, ref.states = pos.copy(), st.copy()
        m.remesh()
        ref = RO.remesh(ref, p)
        s = m.stats()
        added += s["added"]
        removed += s["removed"] + s["out_of_box"]
        assert_same(lib_state(m), ref)
    assert added > 0 and removed > 0


@pytest.mark.parametrize("name", ["2d", "3d"])
def test_laplacian_smoothing_equals_oracle(name):
    """Fluid nodes next to the x = 0 wall are moved to within hchar/3 of it before remeshing:
    the interior ones go to the mean of their neighbours, free-surface ones hchar/3 away from
    the wall (Mesh.cpp:580-760)."""
    m, o, p = _setup(name, laplacianSmoothingBoundaries=True)
    m.triangulate()
    ref = RO.triangulate(o, p)
    lib0 = lib_state(m)
    h = CASES[name]["h"]
    pos = lib0["pos"].copy()
    near = ((lib0["flags"] & BOUND) == 0) & (pos[:, 0] < 1.5 * h)
    pos[near, 0] = 0.2 * h
    m.set_positions_states(pos.T.copy(), None)
    ref.pos = pos.copy()
    m.remesh()
    ref = RO.remesh(ref, p)
    lib = lib_state(m)
    # the line / plane through the facet nodes is solved by Cramer's rule here and by LU in the
    # oracle (QR in the reference): positions agree to rounding
    assert_same(lib, ref, exact=False)
    assert len(lib["pos"]) == len(pos)
    moved = np.abs(lib["pos"] - pos).max(axis=1) > 0
    assert moved[near].sum() >= 3


def test_errors():
    with pytest.raises(pfem_amd.PfemError):
        pfem_amd.Mesher(2, hchar=0.1, boundingBox=[0, 0, 1])
    with pytest.raises(pfem_amd.PfemError):
        pfem_amd.Mesher(4, hchar=0.1)
    m = pfem_amd.Mesher(2, hchar=0.1)
    with pytest.raises(pfem_amd.PfemError):
        m.remesh()   # no triangulation yet
    P = np.array([[0, 1, 2, 3], [0, 0, 0, 0]], dtype=float)   # collinear: no 2D volume
    m.set_nodes(P, np.zeros(4, np.uint8))
    with pytest.raises(pfem_amd.PfemError):
        m.triangulate()
    # alpha so small that no element survives
    m = pfem_amd.Mesher(2, hchar=1e-3, alpha=0.1)
    m.set_nodes(np.random.default_rng(0).random((2, 30)), np.zeros(30, np.uint8))
    with pytest.raises(pfem_amd.PfemError, match="hchar"):
        m.triangulate()
