"""Synthetic unstructured meshes and initial states for the Rocflu hot path.

The reference reads its grids from files written by ``rflupart``/``rfluconv``
(hex2v/tet2v connectivity, node coordinates and per-patch boundary-face
vertex lists; ``modflu/RFLU_ModReadWriteGrid.F90``).  No such files ship with
the reference, so every workload here is generated from a seed.  The arrays
produced below are exactly what such a grid file holds:

* ``xyz``  (nVert, 3) float64 node coordinates,
* ``c2v``  (nCells, 4|8) int32 cell-to-vertex lists (0-based) in Rocflu's
  local vertex order (hex: bottom face 1-4 counter-clockwise seen from +z,
  top face 5-8 above it; tet: negatively oriented so that the reference's
  ``f2vTet`` table, ``modflu/RFLU_ModGrid.F90:92-96``, yields outward normals),
* ``patches``: list of dicts ``{bcType, bcOpt, pout, bf2v (nBF, 3|4)}``;
  boundary-face vertex order is the owning cell's local face order, so the
  face vector of ``libfloflu/FaceVector.F90`` points out of the domain.

Both the HIP library and the CPU oracle consume these same arrays.
"""
import numpy as np

# Boundary-condition kinds handled on the hot path (reference values from
# modfloflu/ModParameters.F90 are not needed at the C boundary; these small
# integers are this library's own enumeration, mirrored in include/rflu_gpu.h).
BC_SLIPWALL = 1
BC_OUTFLOW = 2
BC_NOSLIPWALL = 3
BC_FARFIELD = 4
BC_INJECTION = 5
BC_INFLOW_VELTEMP = 6
BC_INFLOW_TOTANG = 7
BCVALS = 6  # values per patch; their meaning per kind is documented in include/rflu_gpu.h


def _patch(bc, bf2v):
    """Patch dict from a (bcType, bcOpt, pout[, vals]) tuple."""
    t, o, p = bc[:3]
    vals = np.zeros(BCVALS)
    if len(bc) > 3:
        v = np.asarray(bc[3], np.float64)
        vals[:v.size] = v
    return dict(bcType=t, bcOpt=o, pout=p, vals=vals, bf2v=bf2v.astype(np.int64))
BCOPT_SUPERSONIC = 0
BCOPT_SUBSONIC = 1
BCOPT_MIXED = 2

# Local face-to-vertex tables, 0-based transcription of f2vHex / f2vTet
# (modflu/RFLU_ModGrid.F90:92-96).
F2V_HEX = np.array([[0, 3, 2, 1], [0, 1, 5, 4], [1, 2, 6, 5],
                    [2, 3, 7, 6], [0, 4, 7, 3], [4, 5, 6, 7]], dtype=np.int64)
F2V_TET = np.array([[0, 1, 2], [1, 3, 2], [0, 2, 3], [0, 3, 1]], dtype=np.int64)


def _node_grid(nx, ny, nz, lx, ly, lz, jitter, seed):
    """Structured node lattice, jittered by U(-jitter, jitter)*h per axis.

    Interior nodes move along all three axes; a node on a boundary plane
    keeps that plane's coordinate, so boundary nodes move only within their
    plane (edges along their edge; corners stay fixed).
    """
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    i = np.arange(nx + 1, dtype=np.float64)
    j = np.arange(ny + 1, dtype=np.float64)
    k = np.arange(nz + 1, dtype=np.float64)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")  # i fastest
    x = (I * hx).ravel()
    y = (J * hy).ravel()
    z = (K * hz).ravel()
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=(x.size, 3))
        Ii, Ji, Ki = I.ravel(), J.ravel(), K.ravel()
        mx = (Ii > 0) & (Ii < nx)
        my = (Ji > 0) & (Ji < ny)
        mz = (Ki > 0) & (Ki < nz)
        x = x + np.where(mx, d[:, 0] * hx, 0.0)
        y = y + np.where(my, d[:, 1] * hy, 0.0)
        z = z + np.where(mz, d[:, 2] * hz, 0.0)
    return np.ascontiguousarray(np.stack([x, y, z], axis=1))


def _hex_c2v(nx, ny, nz):
    def vid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c2v = np.stack([vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k),
                    vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1),
                    vid(i, j + 1, k + 1)], axis=1)
    return c2v.astype(np.int64), i, j, k


# 5-tet split of a hex (local corners 0..7 in hex order).  Corners with even
# local parity (0,2,5,7) form class A of an even hex, the others class B; for
# an odd hex the classes swap, which makes the split conforming.
_TETS_EVEN = np.array([[0, 2, 5, 7], [1, 0, 2, 5], [3, 0, 2, 7], [4, 0, 5, 7], [6, 2, 5, 7]])
_TETS_ODD = np.array([[1, 3, 4, 6], [0, 1, 3, 4], [2, 1, 3, 6], [5, 1, 4, 6], [7, 3, 4, 6]])


def _orient_tets(tets, xyz):
    """Swap vertices 2,3 where needed so (x2-x1)x(x3-x1).(x4-x1) < 0."""
    p = xyz[tets]
    a = p[:, 1] - p[:, 0]
    b = p[:, 2] - p[:, 0]
    c = p[:, 3] - p[:, 0]
    s = np.einsum("ij,ij->i", np.cross(a, b), c)
    flip = s > 0
    t = tets.copy()
    t[flip, 1], t[flip, 2] = tets[flip, 2], tets[flip, 1]
    return t


def _box_patches(c2v, i, j, k, nx, ny, nz, cell_kind, bc):
    """Boundary faces of the six box sides, one patch per side.

    Patch order: x-, x+, y-, y+, z-, z+.  Within a patch the faces follow the
    cell order.  ``bc`` maps side index -> (bcType, bcOpt, pout[, vals]).
    """
    if cell_kind == 8:
        # hex local faces on each box side: x- -> face 5 (0,4,7,3), x+ -> 2,
        # y- -> 1, y+ -> 3, z- -> 0, z+ -> 5 (0-based rows of F2V_HEX)
        side_face = [4, 2, 1, 3, 0, 5]
        sel = [i == 0, i == nx - 1, j == 0, j == ny - 1, k == 0, k == nz - 1]
        patches = []
        for s in range(6):
            cells = np.nonzero(sel[s])[0]
            bf2v = c2v[cells][:, F2V_HEX[side_face[s]]]
            patches.append(_patch(bc[s], bf2v))
        return patches
    raise ValueError("tet patches are built by tet_box")


def hex_box(nx, ny, nz, lx, ly, lz, jitter=0.0, seed=0, bc=None):
    xyz = _node_grid(nx, ny, nz, lx, ly, lz, jitter, seed)
    c2v, i, j, k = _hex_c2v(nx, ny, nz)
    if bc is None:
        bc = [(BC_OUTFLOW, BCOPT_SUPERSONIC, 0.0)] * 2 + [(BC_SLIPWALL, 0, 0.0)] * 4
    patches = _box_patches(c2v, i, j, k, nx, ny, nz, 8, bc)
    return dict(xyz=xyz, c2v=c2v, nvc=8, patches=patches, dims=(nx, ny, nz), lengths=(lx, ly, lz))


def wedge_hex(nr, nt, nz, rin, rout, angle, xangle, lz, bc=None):
    """A wedge of hexes about the z axis: radii rin..rout (rin > 0), polar
    angles xangle..xangle + angle (radians), z in [0, lz]; node (i, j, k) at
    r = rin + i (rout - rin)/nr, theta = xangle + j angle/nt, z = k lz/nz.
    The (r, theta, z) lattice maps the box's (x, y, z) with its orientation,
    so the hexes are the box's.  Patches in the box's order: r = rin,
    r = rout, theta = xangle, theta = xangle + angle, z = 0, z = lz (slip
    walls by default)."""
    g = _node_grid(nr, nt, nz, 1.0, 1.0, 1.0, 0.0, 0)
    r = rin + g[:, 0] * (rout - rin)
    th = xangle + g[:, 1] * angle
    xyz = np.ascontiguousarray(np.stack([r * np.cos(th), r * np.sin(th), g[:, 2] * lz], axis=1))
    c2v, i, j, k = _hex_c2v(nr, nt, nz)
    if bc is None:
        bc = [(BC_SLIPWALL, 0, 0.0)] * 6
    patches = _box_patches(c2v, i, j, k, nr, nt, nz, 8, bc)
    return dict(xyz=xyz, c2v=c2v, nvc=8, patches=patches, dims=(nr, nt, nz), lengths=(rout - rin, angle, lz))


def tet_box(nx, ny, nz, lx, ly, lz, jitter=0.0, seed=0, bc=None):
    """Hex box split into 5 tets per hex with (i+j+k)-alternating diagonals."""
    xyz = _node_grid(nx, ny, nz, lx, ly, lz, jitter, seed)
    hc2v, i, j, k = _hex_c2v(nx, ny, nz)
    par = (i + j + k) & 1
    loc = np.where(par[:, None, None] == 0, _TETS_EVEN[None], _TETS_ODD[None])  # (nh,5,4)
    tets = np.take_along_axis(hc2v[:, None, :].repeat(5, axis=1), loc, axis=2).reshape(-1, 4)
    tets = _orient_tets(tets, xyz)
    if bc is None:
        bc = [(BC_OUTFLOW, BCOPT_SUPERSONIC, 0.0)] * 2 + [(BC_SLIPWALL, 0, 0.0)] * 4
    # boundary triangles: tet faces whose 3 vertices lie on one box side
    nvx, nvy = nx + 1, ny + 1
    vi = np.arange(xyz.shape[0])
    I = vi % nvx
    J = (vi // nvx) % nvy
    K = vi // (nvx * nvy)
    onside = [I == 0, I == nx, J == 0, J == ny, K == 0, K == nz]
    patches = []
    fv = tets[:, F2V_TET]  # (nt,4,3)
    for s in range(6):
        on = onside[s][fv].all(axis=2)  # (nt,4)
        t_idx, f_idx = np.nonzero(on)
        bf2v = fv[t_idx, f_idx]
        patches.append(_patch(bc[s], bf2v))
    return dict(xyz=xyz, c2v=tets.astype(np.int64), nvc=4, patches=patches,
                dims=(nx, ny, nz), lengths=(lx, ly, lz))


def hex_centroids_approx(mesh):
    """Vertex-average cell centres (only used to place initial states)."""
    return mesh["xyz"][mesh["c2v"]].mean(axis=1)


def normal_shock_state(mach, rho1, p1, gamma):
    """Rankine-Hugoniot post-shock state behind a shock moving at Mach
    ``mach`` into quiescent gas (rho1, p1): returns (rho2, u2, p2)."""
    g = gamma
    c1 = np.sqrt(g * p1 / rho1)
    p2 = p1 * (1.0 + 2.0 * g / (g + 1.0) * (mach * mach - 1.0))
    rho2 = rho1 * ((g + 1.0) * mach * mach) / ((g - 1.0) * mach * mach + 2.0)
    us = mach * c1
    u2 = us * (1.0 - rho1 / rho2)
    return rho2, u2, p2


def prim_to_cons(rho, u, v, w, p, gamma):
    """Conserved (rho, rho u, rho v, rho w, rho E) of a perfect gas."""
    E = p / (rho * (gamma - 1.0)) + 0.5 * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, rho * E], axis=0)


# ---- mixed meshes (tetrahedra, hexahedra, prisms, pyramids) ----
# Local face tables of all four cell kinds (modflu/RFLU_ModGrid.F90:92-102),
# 0-based, -1 for a triangle's 4th vertex.
F2V_BY_NV = {
    4: [[0, 1, 2, -1], [1, 3, 2, -1], [0, 2, 3, -1], [0, 3, 1, -1]],
    5: [[0, 3, 2, 1], [0, 1, 4, -1], [1, 2, 4, -1], [2, 3, 4, -1], [0, 4, 3, -1]],
    6: [[0, 2, 1, -1], [0, 1, 4, 3], [1, 2, 5, 4], [0, 3, 5, 2], [3, 4, 5, -1]],
    8: [list(r) for r in F2V_HEX],
}


def _pad8(cells):
    out = np.full((len(cells), 8), -1, np.int64)
    for q, c in enumerate(cells):
        out[q, :len(c)] = c
    return out


def _mixed_patches(c2v, lat, dims, bc):
    """Boundary faces of a mixed box mesh, one patch per side (x-, x+, y-,
    y+, z-, z+); within a patch the triangles come first, then the
    quadrilaterals (Rocflu's bTri2v / bQuad2v order), each in cell order.
    Vertex order is the owning cell's local face order (outward normals);
    triangles are padded with -1 to 4 entries."""
    nx, ny, nz = dims
    count = {}
    faces = []
    for icl, row in enumerate(c2v):
        nv = int((row >= 0).sum())
        for f in F2V_BY_NV[nv]:
            v = [int(row[q]) for q in f if q >= 0]
            key = tuple(sorted(v))
            count[key] = count.get(key, 0) + 1
            faces.append((key, v))
    sides = [[[], []] for _ in range(6)]
    for key, v in faces:
        if count[key] != 1:
            continue
        L = lat[v]  # lattice (i, j, k) of each vertex; -1 for cell-centre vertices
        if (L < 0).any():
            raise ValueError("boundary face through an interior vertex")
        for s, (ax, val) in enumerate([(0, 0), (0, nx), (1, 0), (1, ny), (2, 0), (2, nz)]):
            if (L[:, ax] == val).all():
                sides[s][len(v) == 4].append(v + [-1] * (4 - len(v)))
                break
        else:
            raise ValueError("open face not on the box surface")
    patches = []
    for s in range(6):
        bf = sides[s][0] + sides[s][1]
        patches.append(_patch(bc[s], np.array(bf, np.int64).reshape(-1, 4)))
    return patches


def mixed_box(nx, ny, nz, lx, ly, lz, layout="tet_prism", split=None, jitter=0.0, seed=0, bc=None):
    """A box of mixed cells on one node lattice, conforming everywhere.

    layout "tet_prism": hexes with i < split are cut into 5 tetrahedra (as in
    tet_box), the others into 2 prisms along x whose triangles carry the
    diagonal the tetrahedra put on the x = split plane.
    layout "prism_hex_pyr": for i < split[0] prisms along z, for
    split[0] <= i < split[1] hexahedra, beyond that each hex is cut into 6
    pyramids around a vertex added at its centre.

    Cells are ordered by kind as Rocflu numbers them (tets, hexes, prisms,
    pyramids); rows are padded with -1 to 8 vertices (nvc = 8), boundary
    faces to 4 (nvf = 4)."""
    xyz = _node_grid(nx, ny, nz, lx, ly, lz, jitter, seed)
    hc2v, i, j, k = _hex_c2v(nx, ny, nz)
    nvx, nvy = nx + 1, ny + 1
    vi = np.arange(xyz.shape[0])
    lat = np.stack([vi % nvx, (vi // nvx) % nvy, vi // (nvx * nvy)], axis=1)
    tets, hexs, pris, pyrs = [], [], [], []
    extra = []
    if layout == "tet_prism":
        ix = nx // 2 if split is None else split
        par = (i + j + k) & 1
        for h in range(hc2v.shape[0]):
            c = hc2v[h]
            if i[h] < ix:
                loc = _TETS_EVEN if par[h] == 0 else _TETS_ODD
                tets.extend(c[loc])
            else:
                # diagonal of the x = ix face of hex (ix-1, j, k) in its tet split
                anti = ((ix - 1 + j[h] + k[h]) & 1) == 0
                # yz square corners A (j,k), B (j+1,k), C (j+1,k+1), D (j,k+1) on the hex's x- and x+ sides
                lo = {"A": c[0], "B": c[3], "C": c[7], "D": c[4]}
                hi = {"A": c[1], "B": c[2], "C": c[6], "D": c[5]}
                tris = [("A", "B", "D"), ("B", "C", "D")] if anti else [("A", "B", "C"), ("A", "C", "D")]
                for t in tris:
                    pris.append([lo[t[0]], lo[t[1]], lo[t[2]], hi[t[0]], hi[t[1]], hi[t[2]]])
        tets = _orient_tets(np.array(tets, np.int64), xyz).tolist() if tets else []
    elif layout == "prism_hex_pyr":
        i1, i2 = (nx // 3, 2 * nx // 3) if split is None else split
        for h in range(hc2v.shape[0]):
            c = hc2v[h]
            if i[h] < i1:
                # prisms along z: the xy square (0,1,2,3) cut along 0-2, counter-clockwise from +z
                pris.append([c[0], c[1], c[2], c[4], c[5], c[6]])
                pris.append([c[0], c[2], c[3], c[4], c[6], c[7]])
            elif i[h] < i2:
                hexs.append(list(c))
            else:
                ctr = xyz.shape[0] + len(extra)
                extra.append(xyz[c].mean(axis=0))
                for f in F2V_HEX:  # base (a, d, c, b) so that face (1, 4, 3, 2) is the hex face (a, b, c, d)
                    a, b, cc, d = c[f]
                    pyrs.append([a, d, cc, b, ctr])
    else:
        raise ValueError(layout)
    if extra:
        xyz = np.ascontiguousarray(np.concatenate([xyz, np.array(extra)]))
        lat = np.concatenate([lat, np.full((len(extra), 3), -1)])
    c2v = _pad8(list(tets) + hexs + pris + pyrs)
    if bc is None:
        bc = [(BC_OUTFLOW, BCOPT_SUPERSONIC, 0.0)] * 2 + [(BC_SLIPWALL, 0, 0.0)] * 4
    patches = _mixed_patches(c2v, lat, (nx, ny, nz), bc)
    return dict(xyz=xyz, c2v=c2v, nvc=8, patches=patches, dims=(nx, ny, nz), lengths=(lx, ly, lz),
                kinds=dict(tet=len(tets), hex=len(hexs), prism=len(pris), pyramid=len(pyrs)))
