"""GMSH ASCII meshes (dgx_mesh_read_gmsh, the data MeshCreator::extract_mesh_data takes from a
GMSH model, S/utils/mesh_creation.cpp:212-330).  The files are written here in both formats the
reader takes (MSH 2.2 and 4.1) for the meshes the structured generators produce, with the
rectangle's physical groups (lines tagged 2..5 named Bottom/Right/Top/Left, the surface 1, as
S/utils/mesh_creation.cpp:248-260 names them); the reader must return the same vertices and
elements and the neighbour table of dgx_mesh_connect_elements (boundary faces carrying the index
of their group).  The host-side reader needs no GPU; the parity run is marked gpu.
"""
import numpy as np
import pytest

import dgx
import refcpu as R


def _boundary_edges(verts, elems, names):
    """Boundary edges (vertex pairs, in element order) grouped by side of the bounding box."""
    x0, y0 = verts.min(axis=0)
    x1, y1 = verts.max(axis=0)
    nv = elems.shape[1]
    count = {}
    for e in elems:
        for f in range(nv):
            a, b = int(e[f]), int(e[(f + 1) % nv])
            count[(min(a, b), max(a, b))] = count.get((min(a, b), max(a, b)), 0) + 1
    sides = {n: [] for n in names}
    for e in elems:
        for f in range(nv):
            a, b = int(e[f]), int(e[(f + 1) % nv])
            if count[(min(a, b), max(a, b))] != 1:
                continue
            pa, pb = verts[a], verts[b]
            if abs(pa[1] - y0) < 1e-12 and abs(pb[1] - y0) < 1e-12:
                side = "Bottom"
            elif abs(pa[0] - x1) < 1e-12 and abs(pb[0] - x1) < 1e-12:
                side = "Right"
            elif abs(pa[1] - y1) < 1e-12 and abs(pb[1] - y1) < 1e-12:
                side = "Top"
            else:
                side = "Left"
            sides[side].append((a, b))
    return sides


NAMES = ["Bottom", "Right", "Top", "Left"]


def _write_v2(path, verts, elems, tag_of=None):
    tag_of = tag_of or {n: 2 + k for k, n in enumerate(NAMES)}
    sides = _boundary_edges(verts, elems, NAMES)
    etype = 2 if elems.shape[1] == 3 else 3
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", str(1 + len(tag_of))]
    lines += [f'1 {t} "{n}"' for n, t in tag_of.items()] + ['2 1 "Domain"', "$EndPhysicalNames"]
    lines += ["$Nodes", str(len(verts))] + [f"{i + 1} {float(x)!r} {float(y)!r} 0" for i, (x, y) in enumerate(verts)] + ["$EndNodes"]
    els = []
    for n, edges in sides.items():
        for a, b in edges:
            if n in tag_of:  # a side without a physical group has no line elements
                els.append(f"1 2 {tag_of[n]} {tag_of[n]} {a + 1} {b + 1}")
    for e in elems:
        els.append(f"{etype} 2 1 1 " + " ".join(str(int(v) + 1) for v in e))
    lines += ["$Elements", str(len(els))] + [f"{k + 1} {s}" for k, s in enumerate(els)] + ["$EndElements"]
    open(path, "w").write("\n".join(lines) + "\n")


def _write_v4(path, verts, elems):
    tag_of = {n: 2 + k for k, n in enumerate(NAMES)}
    sides = _boundary_edges(verts, elems, NAMES)
    etype = 2 if elems.shape[1] == 3 else 3
    x0, y0 = (float(t) for t in verts.min(axis=0))
    x1, y1 = (float(t) for t in verts.max(axis=0))
    L = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", "5"]
    L += [f'1 {t} "{n}"' for n, t in tag_of.items()] + ['2 1 "Domain"', "$EndPhysicalNames"]
    # entities: 4 points, 4 curves (curve k carries physical tag 2 + k), 1 surface
    L += ["$Entities", "4 4 1 0"]
    corners = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    L += [f"{k + 1} {x!r} {y!r} 0 0" for k, (x, y) in enumerate(corners)]
    for k in range(4):
        L.append(f"{k + 1} {x0!r} {y0!r} 0 {x1!r} {y1!r} 0 1 {2 + k} 2 {k + 1} {-(k % 4 + 2) if k < 3 else -1}")
    L += [f"1 {x0!r} {y0!r} 0 {x1!r} {y1!r} 0 1 1 4 1 2 3 4", "$EndEntities"]
    n = len(verts)
    L += ["$Nodes", f"1 {n} 1 {n}", f"2 1 0 {n}"] + [str(i + 1) for i in range(n)]
    L += [f"{float(x)!r} {float(y)!r} 0" for x, y in verts] + ["$EndNodes"]
    blocks, tag = [], 1
    for k, name in enumerate(NAMES):
        edges = sides[name]
        blk = [f"1 {k + 1} 1 {len(edges)}"]
        for a, b in edges:
            blk.append(f"{tag} {a + 1} {b + 1}")
            tag += 1
        blocks.append(blk)
    blk = [f"2 1 {etype} {len(elems)}"]
    for e in elems:
        blk.append(f"{tag} " + " ".join(str(int(v) + 1) for v in e))
        tag += 1
    blocks.append(blk)
    total = tag - 1
    L += ["$Elements", f"{len(blocks)} {total} 1 {total}"] + [s for b in blocks for s in b] + ["$EndElements"]
    open(path, "w").write("\n".join(L) + "\n")


MESHES = {
    "tri": lambda: R.split_quad_mesh(4, 1.0),
    "quad": lambda: R.rect_quad_mesh(0.0, 2.0, 0.0, 1.0, 0.25),
}


@pytest.mark.parametrize("fmt", ["2.2", "4.1"])
@pytest.mark.parametrize("mesh", list(MESHES))
def test_reader_reproduces_the_mesh_and_its_tags(tmp_path, fmt, mesh):
    verts, elems = MESHES[mesh]()
    path = tmp_path / f"m{fmt}.msh"
    (_write_v2 if fmt == "2.2" else _write_v4)(path, verts, elems)
    v, e, nb, names = dgx.read_gmsh(path)
    assert names == NAMES
    assert np.abs(v - verts).max() == 0.0 and np.array_equal(e, elems)
    # the same table as the position-tagged connectivity (Bottom, Right, Top, Left = 0..3)
    assert np.array_equal(nb, dgx.connect(verts, elems, 0))


def test_reader_uses_the_files_group_names(tmp_path):
    verts, elems = MESHES["tri"]()
    # other names and tags: indices follow the tag order; a side without a group gets -1
    tag_of = {"Bottom": 7, "Right": 3, "Top": 5}
    path = tmp_path / "named.msh"
    _write_v2(path, verts, elems, tag_of)
    _, _, nb, names = dgx.read_gmsh(path)
    assert names == ["Right", "Top", "Bottom"]
    want = dgx.connect(verts, elems, 0)
    remap = {0: 2, 1: 0, 2: 1, 3: -1}  # Bottom, Right, Top, Left -> file order
    bnd = want[..., 0] < 0
    assert np.array_equal(nb[~bnd], want[~bnd])
    assert np.array_equal(nb[bnd][:, 1], np.array([remap[int(k)] for k in want[bnd][:, 1]]))


def test_reader_errors(tmp_path):
    with pytest.raises(dgx.DGXError):
        dgx.read_gmsh(tmp_path / "missing.msh")
    bad = tmp_path / "bad.msh"
    bad.write_text("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n")
    with pytest.raises(dgx.DGXError, match="binary"):
        dgx.read_gmsh(bad)
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n$Elements\n0\n$EndElements\n")
    with pytest.raises(dgx.DGXError, match="No 2D elements"):
        dgx.read_gmsh(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("mesh", list(MESHES))
def test_gmsh_mesh_residual_matches_oracle(tmp_path, mesh):
    verts, elems = MESHES[mesh]()
    path = tmp_path / "m.msh"
    _write_v4(path, verts, elems)
    v, e, nb, names = dgx.read_gmsh(path)
    far = [1.0, 0.8, 0.1, 1 / 0.4 + 0.5 * 0.65]
    bcs = {"Bottom": ("far_field", far), "Top": ("slip_wall", [0, 0, 0, 0]), "Left": ("far_field", far)}
    ora = R.Oracle(verts, elems, 2, periodic=0, **({"basis": "dubiner"} if mesh == "tri" else {}))
    for tag, (typ, val) in bcs.items():
        ora.set_bc(tag, typ, val)
    ctx = dgx.Context(v, e, 2, precision="fp64", neighbors=nb, tags=names, bcs=bcs)
    u0 = ora.project(R.vortex_state(ora.volume_points(), xc=0.5, yc=0.5, beta=1.0, length=1.0))
    r = ora.residual(u0)
    assert np.abs(ctx.residual(u0) - r).max() < 1e-11 * np.abs(r).max()
