"""Polyhedral test meshes in Teton's corner arrays: a grid of hexahedra in
which every other hex is cut into six pyramids about its centre. The apex of
a pyramid is a corner with four corner faces, and a pyramid has five
corners, so the mesh exercises the sweep's general corner topology
(snac/SweepUCBxyz.F90:196-215: the outgoing EZ face of a corner with more
than three faces takes the area-weighted mean of the n - 2 incoming FP faces
opposite it).

Test infrastructure only. Corner connectivity follows aux/setTetonZone.F90:
143-191 as the library's hex builder does: each corner's faces are chained
so that the EZ neighbour of corner face i is its counter-clockwise
neighbour on zone face i. Faces are wound so that their right-hand normal
points into the zone; boundary elements are numbered per (zone, face,
corner in face) in zone order. The arrays use the leading dimensions of the
oracle's polyhedral build (oracle.POLY_MAXCF corner faces, POLY_MAXFACES
zone faces)."""
import numpy as np

from oracle import oracle as O

MAXCF, MAXFACES = O.POLY_MAXCF, O.POLY_MAXFACES

# Hex faces over local nodes 0-3 (bottom, counter-clockwise from +z) and 4-7
# (above them), wound inward.
_HEX_FACES = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0)]


def _newell(p):
    n = np.zeros(3)
    for i in range(len(p)):
        a, b = p[i], p[(i + 1) % len(p)]
        n += np.array([(a[1] - b[1]) * (a[2] + b[2]), (a[2] - b[2]) * (a[0] + b[0]),
                       (a[0] - b[0]) * (a[1] + b[1])])
    return n


def _inward(points, nodes, faces):
    """Each face re-wound where needed so that its right-hand normal points
    to the zone's centre."""
    centre = points[list(nodes)].mean(axis=0)
    out = []
    for f in faces:
        p = points[[nodes[i] for i in f]]
        if np.dot(_newell(p), centre - p.mean(axis=0)) < 0:
            f = tuple(reversed(f))
        out.append(tuple(f))
    return out


def hex_pyramid_mesh(nx, ny, nz, jitter=0.12, seed=3, split=lambda i, j, k: (i + j + k) % 2 == 1):
    """Mesh arrays of an nx x ny x nz grid of the unit cube (interior nodes
    moved by up to `jitter` of a cell) whose cells with split(i, j, k) are
    cut into six pyramids; returns the dict TiledProblem.mesh() returns,
    with corner arrays of leading dimension MAXCF and zone-face arrays of
    MAXFACES."""
    rng = np.random.default_rng(seed)
    gx = np.stack(np.meshgrid(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1),
                              np.linspace(0, 1, nz + 1), indexing="ij"), axis=-1)
    inner = np.zeros(gx.shape[:3], bool)
    inner[1:-1, 1:-1, 1:-1] = True
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    gx = gx + inner[..., None] * rng.uniform(-jitter, jitter, gx.shape) * h
    points = list(gx.reshape(-1, 3))

    def nid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    zones = []  # (global node ids, faces over local ids)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                hexn = [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                        nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                        nid(i, j + 1, k + 1)]
                if not split(i, j, k):
                    zones.append((hexn, _HEX_FACES))
                    continue
                apex = len(points)
                points.append(np.mean([points[n] for n in hexn], axis=0))
                for f in _HEX_FACES:
                    nodes = [hexn[v] for v in f] + [apex]
                    faces = [(0, 1, 2, 3)] + [(b, (b + 1) % 4, 4) for b in range(4)]
                    zones.append((nodes, faces))
    points = np.array(points)
    zones = [(nodes, _inward(points, nodes, faces)) for nodes, faces in zones]

    Z = len(zones)
    numCorner = np.array([len(n) for n, _ in zones], np.int32)
    cOffSet = np.concatenate([[0], np.cumsum(numCorner)[:-1]]).astype(np.int32)
    Cn = int(numCorner.sum())
    zoneFaces = np.array([len(f) for _, f in zones], np.int32)
    assert zoneFaces.max() <= MAXFACES and numCorner.max() <= O.POLY_MAXCORNER
    px = np.zeros((Cn, 3))
    for z, (nodes, _) in enumerate(zones):
        px[cOffSet[z]:cOffSet[z] + len(nodes)] = points[nodes]

    # Face matching by node set.
    zoneOpp = np.zeros((Z, MAXFACES), np.int32)
    faceOpp = np.zeros((Z, MAXFACES), np.int32)
    seen = {}
    for z, (nodes, faces) in enumerate(zones):
        for f, fc in enumerate(faces):
            key = tuple(sorted(nodes[v] for v in fc))
            if key in seen:
                zn, fn = seen.pop(key)
                zoneOpp[z, f], faceOpp[z, f] = zn + 1, fn + 1
                zoneOpp[zn, fn], faceOpp[zn, fn] = z + 1, f + 1
            else:
                seen[key] = (z, f)
                zoneOpp[z, f] = -1

    # Corner across each (zone, face, corner in face): the neighbour zone's
    # corner at the same node, or a new boundary element.
    nbelem = 0
    opp = {}
    for z, (nodes, faces) in enumerate(zones):
        for f, fc in enumerate(faces):
            zn = zoneOpp[z, f]
            for v in fc:
                if zn < 0:
                    nbelem += 1
                    opp[z, f, v] = Cn + nbelem
                else:
                    nn = zones[zn - 1][0]
                    opp[z, f, v] = int(cOffSet[zn - 1]) + nn.index(nodes[v]) + 1

    nCFaces = np.zeros(Cn, np.int32)
    cFP = np.zeros((Cn, MAXCF), np.int32)
    cEZ = np.zeros((Cn, MAXCF), np.int32)
    CToFace = np.zeros((Cn, MAXCF), np.int32)
    for z, (nodes, faces) in enumerate(zones):
        ent = [[] for _ in nodes]  # per local corner: (cw, ccw, opp, face)
        for f, fc in enumerate(faces):
            n = len(fc)
            for i, v in enumerate(fc):
                ent[v].append((fc[(i + 1) % n] + 1, fc[(i + n - 1) % n] + 1, opp[z, f, v], f + 1))
        for v, e in enumerate(ent):
            cc = int(cOffSet[z]) + v
            n = len(e)
            assert n <= MAXCF
            nCFaces[cc] = n
            cEZ[cc, 0], cFP[cc, 0], CToFace[cc, 0] = e[0][1], e[0][2], e[0][3]
            last = e[0][0]
            for i in range(1, n):
                nxt = [t for t in e[1:] if t[1] == last]
                assert len(nxt) == 1, "corner faces do not close"
                cEZ[cc, i], cFP[cc, i], CToFace[cc, i] = nxt[0][1], nxt[0][2], nxt[0][3]
                last = nxt[0][0]
    return dict(numCorner=numCorner, cOffSet=cOffSet, nCFaces=nCFaces, cFP=cFP, cEZ=cEZ,
                CToFace=CToFace, zoneOpp=zoneOpp, faceOpp=faceOpp, px=px, zoneFaces=zoneFaces,
                nbelem=nbelem)


def widen(mesh):
    """Hex mesh arrays (corner faces 3, zone faces 6) laid out with the
    polyhedral build's leading dimensions."""
    out = dict(mesh)
    for k in ("cFP", "cEZ", "CToFace"):
        a = np.zeros((mesh[k].shape[0], MAXCF), np.int32)
        a[:, :3] = mesh[k]
        out[k] = a
    for k in ("zoneOpp", "faceOpp"):
        a = np.zeros((mesh[k].shape[0], MAXFACES), np.int32)
        a[:, :6] = mesh[k]
        out[k] = a
    return out


class PolyProblem:
    """Mesh, geometry (the oracle's), quadrature and schedule (the oracle's)
    of a polyhedral mesh, in the form the drop-in wrapper reads."""

    def __init__(self, mesh, quadrature, par_tol=0.0):
        self._m = mesh
        self.nzones = len(mesh["numCorner"])
        self.ncorner = len(mesh["nCFaces"])
        self.nbelem = int(mesh["nbelem"]) if "nbelem" in mesh else max(0, int(mesh["cFP"].max()) - self.ncorner)
        A_fp, A_ez, V = O.geometry(mesh)
        self._g = dict(A_fp=A_fp, A_ez=A_ez, Volume=V)
        self._q = quadrature
        self.nangles = len(quadrature["weight"])
        self._s = O.schedule(mesh, A_fp, A_ez, quadrature["omega"], par_tol)

    def mesh(self):
        return self._m

    def geometry(self):
        return self._g

    def quadrature(self):
        return self._q

    def schedule(self):
        return self._s
