"""Gmsh .msh input (Gmsh3DSingle): the reader and its connectivity builder.

No .msh file ships with the reference, so the files are written here from
meshes whose connectivity is known: the uniform box (its EtoE/EtoF/EtoR/EtoB
come from the hex-split neighbour table), written as MSH 2.2 and 4.1 with the
periodic node pairs, must read back to the same arrays; a relabelled and
vertex-reordered copy with moved vertices must give face node pairs that
coincide in space and a free stream that stays free."""
import os

import numpy as np
import pytest

from exadglm_amd.gmsh import Gmsh3DSingle, read_msh
from exadglm_amd.mesh import Mesh3D, Uniform3DSingle
from exadglm_amd.reference import Operator3DFace, Operator3DRef, TetraRef


def _periodic_pairs(V, box, axes):
    """slave -> master node indices for the periodic axes of a box mesh."""
    lo = np.array(box[0::2])
    pairs = {}
    key = {tuple(np.round(v, 12)): i for i, v in enumerate(V)}
    for i, v in enumerate(V):
        for a in axes:
            if abs(v[a] - box[2 * a + 1]) < 1e-12:
                w = v.copy()
                w[a] = lo[a]
                pairs[i] = key[tuple(np.round(w, 12))]
    return pairs


def _boundary_tris(elem):
    """(3 vertex indices, tag) of the faces with no partner."""
    fm = ((1, 0, 2), (0, 1, 3), (1, 2, 3), (2, 0, 3))
    out = []
    for e, f in zip(*np.nonzero(elem.EtoE < 0)):
        out.append(([int(elem.EtoV[e, v]) for v in fm[f]], int(elem.EtoB[e, f])))
    return out


def write_msh(path, V, EtoV, tris=(), pairs=None, version="2.2", tags=None):
    """Write an ASCII .msh file (node tags = tags[i] or i + 1)."""
    tags = np.arange(1, len(V) + 1) if tags is None else tags
    pairs = pairs or {}
    L = ["$MeshFormat", f"{version} 0 8", "$EndMeshFormat"]
    if version == "2.2":
        L += ["$Nodes", str(len(V))] + [f"{tags[i]} {float(x)!r} {float(y)!r} {float(z)!r}" for i, (x, y, z) in enumerate(V)]
        L += ["$EndNodes", "$Elements", str(len(tris) + len(EtoV))]
        k = 1
        for t, ph in tris:
            L.append(f"{k} 2 2 {ph} 1 " + " ".join(str(tags[v]) for v in t))
            k += 1
        for tet in EtoV:
            L.append(f"{k} 4 2 0 1 " + " ".join(str(tags[v]) for v in tet))
            k += 1
        L.append("$EndElements")
        if pairs:
            L += ["$Periodic", "1", "2 2 1", str(len(pairs))] + [f"{tags[s]} {tags[m]}" for s, m in pairs.items()]
            L.append("$EndPeriodic")
    else:
        phys = sorted({ph for _, ph in tris})
        # one surface entity per physical tag, one volume entity
        L += ["$Entities", f"0 0 {len(phys)} 1"]
        L += [f"{i + 1} 0 0 0 1 1 1 1 {ph} 0" for i, ph in enumerate(phys)]
        L += ["1 0 0 0 1 1 1 0 0", "$EndEntities"]
        L += ["$Nodes", f"1 {len(V)} 1 {len(V)}", f"3 1 0 {len(V)}"]
        L += [str(t) for t in tags] + [f"{float(x)!r} {float(y)!r} {float(z)!r}" for x, y, z in V] + ["$EndNodes"]
        blocks = []
        k = 1
        for i, ph in enumerate(phys):
            rows = [t for t, p in tris if p == ph]
            blocks.append(f"2 {i + 1} 2 {len(rows)}")
            for t in rows:
                blocks.append(f"{k} " + " ".join(str(tags[v]) for v in t))
                k += 1
        blocks.append(f"3 1 4 {len(EtoV)}")
        for tet in EtoV:
            blocks.append(f"{k} " + " ".join(str(tags[v]) for v in tet))
            k += 1
        L += ["$Elements", f"{len(phys) + 1} {k - 1} 1 {k - 1}"] + blocks + ["$EndElements"]
        if pairs:
            L += ["$Periodic", "1", "2 2 1", "0", str(len(pairs))] + [f"{tags[s]} {tags[m]}" for s, m in pairs.items()]
            L.append("$EndPeriodic")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


@pytest.mark.parametrize("version", ["2.2", "4.1"])
@pytest.mark.parametrize("bc", [{"x": "pbc", "y": "pbc", "z": "pbc"}, {"x": "bc2", "y": "pbc", "z": "bc3"}])
def test_uniform_box_round_trip(tmp_path, version, bc):
    box = (0.0, 1.0, -1.0, 2.0, 0.5, 1.5)
    u = Uniform3DSingle(4, 5, 3, *box, bc=bc)
    axes = [a for a, n in enumerate("xyz") if bc[n] == "pbc"]
    path = os.path.join(tmp_path, "box.msh")
    write_msh(path, u.VXYZ, u.EtoV, _boundary_tris(u), _periodic_pairs(u.VXYZ, box, axes), version)
    g = Gmsh3DSingle(path)
    assert g.Nelem == u.Nelem
    np.testing.assert_array_equal(g.VXYZ, u.VXYZ)
    np.testing.assert_array_equal(g.EtoV, u.EtoV)
    np.testing.assert_array_equal(g.EtoE, u.EtoE)
    np.testing.assert_array_equal(g.EtoB, u.EtoB)
    # EtoF and EtoR of a face without partner are not read (its vmapP is its vmapM)
    inner = u.EtoE >= 0
    np.testing.assert_array_equal(g.EtoF[inner], u.EtoF[inner])
    np.testing.assert_array_equal(g.EtoR[inner], u.EtoR[inner])


def test_relabelled_moved_mesh(tmp_path):
    """Node tags shuffled, elements shuffled, each tet's vertices cyclically
    shifted or mirrored (orientation restored by the reader), interior
    vertices moved: every face node pair coincides modulo the period, and a
    uniform Euler state has zero RHS."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import pyoracle as O
    from exadglm_amd.problem import DATAIN_NAMES, Problem

    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    u = Uniform3DSingle(5, 5, 5, *box)
    rng = np.random.default_rng(4)
    V = u.VXYZ.copy()
    interior = np.all((V > 1e-12) & (V < 1 - 1e-12), axis=1)
    V[interior] += rng.uniform(-0.04, 0.04, size=(interior.sum(), 3))
    perm = rng.permutation(len(V))                   # node i is written with tag perm[i] + 1
    E = u.EtoV[rng.permutation(u.Nelem)].copy()
    for row in E:
        k = rng.integers(0, 4)
        if k == 1:
            row[:] = row[[1, 2, 0, 3]]
        elif k == 2:
            row[:] = row[[0, 2, 1, 3]]
        elif k == 3:
            row[:] = row[[3, 1, 2, 0]]
    path = os.path.join(tmp_path, "moved.msh")
    write_msh(path, V, E, (), _periodic_pairs(u.VXYZ, box, [0, 1, 2]), "4.1", tags=perm + 1)
    g = Gmsh3DSingle(path)
    N = 3
    tet = TetraRef(N)
    op_face = Operator3DFace(Operator3DRef(tet))
    m = Mesh3D(op_face, g)
    assert np.all(m.J > 0)
    P = np.stack([m.PX.ravel(), m.PY.ravel(), m.PZ.ravel()], axis=1)
    d = P[m.vmapM.ravel()] - P[m.vmapP.ravel()]
    d -= np.round(d)                                  # period 1 on every axis
    assert np.abs(d).max() < 1e-12
    np.testing.assert_array_equal(m.EtoE[m.EtoE[np.arange(g.Nelem)[:, None], np.arange(4)[None]],
                                         m.EtoF], np.arange(g.Nelem)[:, None] + 0 * m.EtoF)
    # free stream: rho=1, v=(0.3,-0.2,0.1), p=1 has zero RHS on any mesh
    c = np.ascontiguousarray
    arr = {"EtoB": c(m.EtoB, np.int32), "vmapP": c(m.vmapP, np.int32), "vmapF": c(m.vmapF, np.int32),
           "Fmask": c(op_face.Fmask, np.int32), "WDr": c(m.op_ref.WDr), "WDs": c(m.op_ref.WDs),
           "WDt": c(m.op_ref.WDt), "LIFT": c(op_face.LIFT), "Fscale": c(m.Fscale), "rx": c(m.rx), "ry": c(m.ry),
           "rz": c(m.rz), "sx": c(m.sx), "sy": c(m.sy), "sz": c(m.sz), "tx": c(m.tx), "ty": c(m.ty), "tz": c(m.tz),
           "nx": c(m.NX), "ny": c(m.NY), "nz": c(m.NZ), "px": c(m.PX), "py": c(m.PY), "pz": c(m.PZ)}
    assert sorted(arr) == sorted(DATAIN_NAMES)
    pb = Problem(N=N, K=g.Nelem, Np=tet.Np, Nfp=tet.Nfp, tet=tet, op_ref=m.op_ref, op_face=op_face, mesh=m, arrays=arr)
    KNp = g.Nelem * tet.Np
    v = (0.3, -0.2, 0.1)
    uu = np.zeros((5, KNp))
    uu[0] = 1.0
    uu[1:4] = np.array(v)[:, None]
    uu[4] = 1.0 / 0.4 + 0.5 * sum(x * x for x in v)
    r, _, _ = O.Oracle(pb, "euler").rhs(uu)
    assert np.abs(r).max() < 1e-11


def test_reader_errors(tmp_path):
    p = os.path.join(tmp_path, "bad.msh")
    with open(p, "w") as fh:
        fh.write("$MeshFormat\n3.0 0 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError):
        read_msh(p)
    with open(p, "w") as fh:
        fh.write("$MeshFormat\n2.2 1 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError):
        read_msh(p)
    with open(p, "w") as fh:
        fh.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n$Elements\n0\n$EndElements\n")
    with pytest.raises(ValueError):
        read_msh(p)
