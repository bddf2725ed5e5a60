"""gmsh MSH 4.1 ASCII reader: the node cloud the reference builds from a .msh file.

``Mesh::loadFromFile`` (``srcs/mesh/Mesh.cpp:762-917``) opens the file with the gmsh SDK, takes
the mesh dimension from the element types (``Mesh::computeMeshDim``, ``Mesh.cpp:209-235``: the
largest of 2 and 3 that has elements; two element types of one dimension are an error), and
builds its node list from the physical groups, dropping the elements (the alpha-shape
triangulation that follows rebuilds them):

1. every physical group of dimension dim-1 except "FreeSurface" and "Fluid", in the order
   ``gmsh::model::getPhysicalGroups`` lists them (by tag): its nodes become boundary nodes
   tagged with the group's name;
2. every physical group of dimension dim, in the same order: its nodes become interior
   nodes tagged with the group's name;

a node whose position equals one already in the list is skipped (``Node::operator==``), and
the tag names are numbered in order of first appearance.  ``getNodesForPhysicalGroup`` returns
the nodes of the group's elements once each, in increasing node tag, which this reader does
too.  The gmsh SDK is not part of this package; this module parses the file format itself.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["MshError", "NodeCloud", "read_msh", "load_node_cloud"]

# gmsh element type -> (dimension, nodes per element) for the types a PFEM mesh file holds
_ELEM = {15: (0, 1), 1: (1, 2), 8: (1, 3), 2: (2, 3), 3: (2, 4), 9: (2, 6), 4: (3, 4), 5: (3, 8), 6: (3, 6),
         7: (3, 5), 11: (3, 10)}


class MshError(RuntimeError):
    pass


@dataclass
class MshFile:
    physical_names: dict = field(default_factory=dict)    # (dim, tag) -> name
    entity_phys: dict = field(default_factory=dict)       # (dim, entity tag) -> [physical tags]
    node_coords: dict = field(default_factory=dict)       # node tag -> (x, y, z)
    elements: list = field(default_factory=list)          # (dim, entity tag, type, [[node tags]])


def _sections(lines):
    i = 0
    while i < len(lines):
        line = lines[i].strip()
        if line.startswith("$") and not line.startswith("$End"):
            name = line[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != "$End" + name:
                j += 1
            if j == len(lines):
                raise MshError(f"section ${name} is not closed")
            yield name, [ln for ln in lines[i + 1:j]]
            i = j + 1
        else:
            i += 1


def read_msh(path: str) -> MshFile:
    with open(path) as f:
        lines = f.read().splitlines()
    msh = MshFile()
    seen = set()
    for name, body in _sections(lines):
        seen.add(name)
        tok = " ".join(body).split()
        if name == "MeshFormat":
            version, ftype = float(tok[0]), int(tok[1])
            if not (4.0 <= version < 5.0) or ftype != 0:
                raise MshError(f"only MSH 4.x ASCII is read (got version {tok[0]}, file type {ftype})")
        elif name == "PhysicalNames":
            for ln in body[1:]:
                parts = ln.split(None, 2)
                if len(parts) == 3:
                    msh.physical_names[(int(parts[0]), int(parts[1]))] = parts[2].strip().strip('"')
        elif name == "Entities":
            it = iter(tok)
            counts = [int(next(it)) for _ in range(4)]
            for dim, cnt in enumerate(counts):
                for _ in range(cnt):
                    tag = int(next(it))
                    nbox = 3 if dim == 0 else 6
                    for _ in range(nbox):
                        next(it)
                    nphys = int(next(it))
                    msh.entity_phys[(dim, tag)] = [int(next(it)) for _ in range(nphys)]
                    if dim > 0:
                        nb = int(next(it))
                        for _ in range(nb):
                            next(it)
        elif name == "Nodes":
            it = iter(tok)
            nblocks = int(next(it))
            next(it), next(it), next(it)  # numNodes, minTag, maxTag
            for _ in range(nblocks):
                next(it), next(it)  # entity dim, entity tag
                parametric = int(next(it))
                n = int(next(it))
                tags = [int(next(it)) for _ in range(n)]
                for t in tags:
                    xyz = (float(next(it)), float(next(it)), float(next(it)))
                    if parametric:
                        raise MshError("parametric node coordinates are not read")
                    msh.node_coords[t] = xyz
        elif name == "Elements":
            it = iter(tok)
            nblocks = int(next(it))
            next(it), next(it), next(it)
            for _ in range(nblocks):
                edim, etag, etype, n = int(next(it)), int(next(it)), int(next(it)), int(next(it))
                if etype not in _ELEM:
                    raise MshError(f"unsupported element type {etype}")
                npe = _ELEM[etype][1]
                conn = []
                for _ in range(n):
                    next(it)  # element tag
                    conn.append([int(next(it)) for _ in range(npe)])
                msh.elements.append((edim, etag, etype, conn))
    for need in ("MeshFormat", "Nodes", "Elements"):
        if need not in seen:
            raise MshError(f"missing ${need} section")
    return msh


def mesh_dim(msh: MshFile) -> int:
    """Mesh::computeMeshDim (Mesh.cpp:209-235)."""
    dim = -1
    for d in (2, 3):
        types = {t for (ed, _e, t, _c) in msh.elements if ed == d}
        if len(types) == 1:
            dim = d
        elif len(types) > 1:
            raise MshError("do not use hybrid meshes for PFEM simulations!")
    if dim == -1:
        raise MshError("there is no suitable elements in the .msh file!")
    return dim


def nodes_for_physical_group(msh: MshFile, dim: int, tag: int):
    """gmsh::model::mesh::getNodesForPhysicalGroup: the nodes of the group's elements, each once,
    by increasing node tag."""
    ents = {e for (d, e), phys in msh.entity_phys.items() if d == dim and tag in phys}
    nodes = set()
    for edim, etag, _t, conn in msh.elements:
        if edim == dim and etag in ents:
            for c in conn:
                nodes.update(c)
    return sorted(nodes)


@dataclass
class NodeCloud:
    dim: int
    pos: np.ndarray       # dim x N, field-blocked (the library's layout)
    bound: np.ndarray     # N bools (Node::m_isBound)
    tag: np.ndarray       # N ints, index into tag_names (Node::m_tag)
    tag_names: list       # Mesh::m_tagNames

    @property
    def n_nodes(self):
        return self.pos.shape[1]


def load_node_cloud(path: str, bounding_box=None) -> NodeCloud:
    """The node list Mesh::loadFromFile builds (Mesh.cpp:762-917) before triangulating.
    bounding_box (2 dim values) is checked for size as the reference does."""
    msh = read_msh(path)
    dim = mesh_dim(msh)
    if bounding_box is not None and len(bounding_box) != 2 * dim:
        raise MshError(f"Invalid bounding box size: {len(bounding_box)}")
    groups_ld = sorted(t for (d, t) in msh.physical_names if d == dim - 1)
    groups_hd = sorted(t for (d, t) in msh.physical_names if d == dim)
    positions, bound, tags, names = [], [], [], []
    seen = set()

    def add(group_dim, gtag, is_bound):
        name = msh.physical_names[(group_dim, gtag)]
        for nt in nodes_for_physical_group(msh, group_dim, gtag):
            xyz = msh.node_coords[nt][:dim]
            if xyz in seen:  # the same position is already in the list (Node::operator==)
                continue
            seen.add(xyz)
            if name not in names:
                names.append(name)
            positions.append(xyz)
            bound.append(is_bound)
            tags.append(names.index(name))

    for g in groups_ld:
        if msh.physical_names[(dim - 1, g)] not in ("FreeSurface", "Fluid"):
            add(dim - 1, g, True)
    for g in groups_hd:
        add(dim, g, False)
    if not positions:
        raise MshError("no nodes loaded! Did you add the physical groups in the .geo file?")
    pos = np.array(positions, dtype=np.float64).T.copy()
    return NodeCloud(dim, pos, np.array(bound), np.array(tags, dtype=np.int32), names)
