"""Gmsh mesh input: `Gmsh3DSingle`, the mesh source the reference's launcher
imports for mesh_info = {'type': 'gmsh', 'file': path}
(dim3/src/python/dglm_3d_launcher.py:362-373; the `gmsh_3d_single` module
itself is not shipped upstream).

Reads ASCII MSH 2.2 and 4.1 files: 4-node tetrahedra (element type 4) are the
elements, 3-node triangles (type 2) give the physical tags of boundary faces,
and `$Periodic` node correspondences identify the periodic copies of nodes.
The result has the attributes of `Uniform3DSingle` (VXYZ, EtoV, EtoE, EtoF,
EtoR, EtoB, Nelem, Nface), so `Mesh3D` builds the node coordinates, metrics,
normals and vmaps from it unchanged.

Connectivity follows the uniform mesh's conventions (mesh.cpp,
uniform_3d_single_core.cpp:70-277):
  * local faces of a tet: 0 = (v1,v0,v2), 1 = (v0,v1,v3), 2 = (v1,v2,v3),
    3 = (v2,v0,v3) -- the N = 1 Fmask rows;
  * EtoE/EtoF: the element and local face sharing the face's three vertices
    (periodic copies identified);
  * EtoR: the rotation r for which vertex EtoV[e][Fmask1[f][j]] is vertex
    EtoV[EtoE][FmaskP1[EtoF][r][j]] for j = 0, 1, 2 -- the N = 1 face masks,
    so that vmapP = EtoE*Np + FmaskP[EtoF][EtoR] pairs coinciding nodes for
    every order;
  * faces with no partner: EtoE = -99, EtoB = the physical tag of the
    boundary triangle on it (0 without one); interior faces EtoB = -99.
Tetrahedra are reordered to positive orientation (v1 and v2 swapped when the
signed volume is negative).
"""
import numpy as np

# N = 1 face masks (reference-element vertices on each face, in the order the
# face node lists use) and the partner orders per rotation; the N = 1 tables
# of Operator3DFace (the general-N tables reduce to these).
FMASK1 = ((1, 0, 2), (0, 1, 3), (1, 2, 3), (2, 0, 3))
FMASKP1 = (((0, 1, 2), (1, 2, 0), (2, 0, 1)),
           ((1, 0, 3), (0, 3, 1), (3, 1, 0)),
           ((2, 1, 3), (1, 3, 2), (3, 2, 1)),
           ((0, 2, 3), (2, 3, 0), (3, 0, 2)))


class MshData:
    """What a .msh file says about a tetrahedral mesh (node tags as in the file)."""

    def __init__(self):
        self.nodes = {}          # node tag -> (x, y, z)
        self.tets = []           # 4 node tags each
        self.tris = []           # (3 node tags, physical tag)
        self.periodic = {}       # slave node tag -> master node tag


def _sections(lines):
    """Map section name -> list of lines between $Name and $EndName."""
    out, i = {}, 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("$") and not ln.startswith("$End"):
            name = ln[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != "$End" + name:
                j += 1
            if j == len(lines):
                raise ValueError(f"msh: section ${name} has no $End{name}")
            out[name] = [x for x in lines[i + 1:j]]
            i = j + 1
        else:
            i += 1
    return out


def _read_v2(sec, d):
    body = sec["Nodes"]
    n = int(body[0])
    for ln in body[1:1 + n]:
        f = ln.split()
        d.nodes[int(f[0])] = (float(f[1]), float(f[2]), float(f[3]))
    body = sec["Elements"]
    m = int(body[0])
    for ln in body[1:1 + m]:
        f = [int(x) for x in ln.split()]
        etype, ntags = f[1], f[2]
        tags, conn = f[3:3 + ntags], f[3 + ntags:]
        if etype == 4:
            d.tets.append(conn[:4])
        elif etype == 2:
            d.tris.append((conn[:3], tags[0] if ntags else 0))
    if "Periodic" in sec:
        body = sec["Periodic"]
        nlinks, k = int(body[0]), 1
        for _ in range(nlinks):
            k += 1                                   # dim slaveTag masterTag
            if body[k].split()[0].lower() == "affine":
                k += 1
            nn = int(body[k])
            k += 1
            for ln in body[k:k + nn]:
                s, m_ = ln.split()[:2]
                d.periodic[int(s)] = int(m_)
            k += nn


def _read_v4(sec, d):
    # physical tag of each surface entity
    surf_phys = {}
    if "Entities" in sec:
        body = sec["Entities"]
        npt, ncu, nsu, nvo = (int(x) for x in body[0].split()[:4])
        k = 1
        for _ in range(npt):
            k += 1
        for _ in range(ncu):
            k += 1
        for _ in range(nsu):
            f = body[k].split()
            nph = int(f[7])
            surf_phys[int(f[0])] = int(f[8]) if nph > 0 else 0
            k += 1
    body = sec["Nodes"]
    nblocks = int(body[0].split()[0])
    k = 1
    for _ in range(nblocks):
        edim, etag, param, nin = (int(x) for x in body[k].split()[:4])
        k += 1
        tags = [int(body[k + i]) for i in range(nin)]
        k += nin
        for i in range(nin):
            f = body[k + i].split()
            d.nodes[tags[i]] = (float(f[0]), float(f[1]), float(f[2]))
        k += nin
    body = sec["Elements"]
    nblocks = int(body[0].split()[0])
    k = 1
    for _ in range(nblocks):
        edim, etag, etype, nin = (int(x) for x in body[k].split()[:4])
        k += 1
        for ln in body[k:k + nin]:
            f = [int(x) for x in ln.split()]
            if etype == 4:
                d.tets.append(f[1:5])
            elif etype == 2:
                d.tris.append((f[1:4], surf_phys.get(etag, 0)))
        k += nin
    if "Periodic" in sec:
        body = sec["Periodic"]
        nlinks, k = int(body[0]), 1
        for _ in range(nlinks):
            k += 1                                   # entityDim entityTag entityTagMaster
            naff = int(body[k].split()[0])
            k += 1 if naff == 0 or len(body[k].split()) > 1 else 2
            nn = int(body[k])
            k += 1
            for ln in body[k:k + nn]:
                s, m_ = ln.split()[:2]
                d.periodic[int(s)] = int(m_)
            k += nn


def read_msh(path):
    """Parse an ASCII MSH 2.2 or 4.1 file into MshData."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    sec = _sections(lines)
    if "MeshFormat" not in sec:
        raise ValueError("msh: no $MeshFormat section")
    fmt = sec["MeshFormat"][0].split()
    version, ftype = float(fmt[0]), int(fmt[1])
    if ftype != 0:
        raise ValueError("msh: only ASCII files are read")
    d = MshData()
    if 2.0 <= version < 3.0:
        _read_v2(sec, d)
    elif 4.0 <= version < 5.0:
        _read_v4(sec, d)
    else:
        raise ValueError(f"msh: format version {version} is not read (2.2 and 4.1 are)")
    if not d.tets:
        raise ValueError("msh: no 4-node tetrahedra")
    return d


def connect(EtoV, vid, btag=None):
    """EtoE, EtoF, EtoR, EtoB of tetrahedra EtoV [K,4] whose vertices are
    identified by vid (vertex -> canonical id; periodic copies share one).
    btag: dict frozenset(canonical vertex ids) -> physical tag of boundary
    faces."""
    K = EtoV.shape[0]
    cid = vid[EtoV]                                   # [K, 4] canonical ids
    EtoE = np.full((K, 4), -99, np.int32)
    EtoF = np.zeros((K, 4), np.int32)
    EtoR = np.zeros((K, 4), np.int32)
    EtoB = np.full((K, 4), -99, np.int32)
    faces = {}
    for e in range(K):
        for f in range(4):
            key = frozenset(int(cid[e, v]) for v in FMASK1[f])
            if len(key) != 3:
                raise ValueError(f"msh: element {e} face {f} has repeated vertices")
            faces.setdefault(key, []).append((e, f))
    for key, side in faces.items():
        if len(side) > 2:
            raise ValueError("msh: a face is shared by more than two tetrahedra")
        if len(side) == 1:
            e, f = side[0]
            EtoE[e, f] = -99
            EtoF[e, f] = f
            EtoB[e, f] = (btag or {}).get(key, 0)
            continue
        for (e, f), (nb, nf) in (side, side[::-1]):
            mine = [int(cid[e, v]) for v in FMASK1[f]]
            for r in range(3):
                if mine == [int(cid[nb, v]) for v in FMASKP1[nf][r]]:
                    break
            else:
                raise ValueError(f"msh: faces ({e},{f}) and ({nb},{nf}) do not match under any rotation")
            EtoE[e, f], EtoF[e, f], EtoR[e, f] = nb, nf, r
    return EtoE, EtoF, EtoR, EtoB


class Gmsh3DSingle:
    """Tetrahedral mesh read from a Gmsh .msh file (single partition)."""

    def __init__(self, path, mod_suffix="cpp"):
        d = read_msh(path)
        tags = sorted(d.nodes)
        index = {t: i for i, t in enumerate(tags)}
        self.VXYZ = np.array([d.nodes[t] for t in tags], dtype=np.float64)
        EtoV = np.array([[index[t] for t in tet] for tet in d.tets], dtype=np.int32)
        # positive orientation: det[v1-v0, v2-v0, v3-v0] > 0
        x = self.VXYZ[EtoV]
        det = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])
        if np.any(det == 0.0):
            raise ValueError("msh: degenerate tetrahedron")
        neg = det < 0
        EtoV[neg, 1], EtoV[neg, 2] = EtoV[neg, 2].copy(), EtoV[neg, 1].copy()
        # canonical vertex ids: follow periodic slave -> master chains
        vid = np.arange(len(tags), dtype=np.int64)
        for s, m in d.periodic.items():
            if s in index and m in index:
                vid[index[s]] = index[m]
        for _ in range(3):                             # corner nodes can chain through two links
            vid = vid[vid]
        btag = {frozenset(int(vid[index[t]]) for t in tri): ph for tri, ph in d.tris}
        self.EtoV = np.ascontiguousarray(EtoV)
        self.EtoE, self.EtoF, self.EtoR, self.EtoB = connect(self.EtoV, vid, btag)
        self.Nelem = EtoV.shape[0]
        self.Nface = 4
