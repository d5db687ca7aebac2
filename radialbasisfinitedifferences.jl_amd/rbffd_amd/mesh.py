"""CGNS mesh import, boundary normals and ghost nodes: the host-side mesh
path of the reference (src/processmesh.jl:1-218, src/extractcoordinates.jl,
src/extractelements.jl, src/calculatenormal.jl, src/genghostnodes.jl).

Node sets follow the reference: the nodes are the centroids of the mesh's
triangles (interior) and the midpoints of its boundary line elements, in CGNS
element order, and every boundary node has a ghost node at distance h along
its outward normal.  Indices here are 0-based.  Nearest-neighbour searches are
exact (the reference uses an HNSW graph, an approximate search that finds the
same single nearest node on these meshes).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

from .hdf5 import H5File


def extractcoordinates(fid: H5File, field_x: str, field_y: str):
    """(P, 2) mesh vertex coordinates (src/extractcoordinates.jl:1-13)."""
    x = fid.read(field_x + "/ data").astype(np.float64).ravel()
    y = fid.read(field_y + "/ data").astype(np.float64).ravel()
    return np.ascontiguousarray(np.stack([x, y], 1))


def extractelements(fid: H5File, X, field: str, conn: int):
    """Element centroids and connectivity (src/extractelements.jl:1-19).
    Connectivity is returned 0-based, (elements, conn); the centroid is the
    mean of the element's vertices, summed in vertex order."""
    c = fid.read(field).astype(np.int64).ravel().reshape(-1, conn) - 1
    pts = X[c]                       # (elements, conn, 2)
    s = pts[:, 0, :].copy()
    for j in range(1, conn):
        s = s + pts[:, j, :]
    return s / conn, c


def calculatenormal(X, elements):
    """Unit normals (-dy, dx)/len and tangents (dx, dy)/len of line elements
    (src/calculatenormal.jl:1-16).  Returns (normals, tangents), each (E, 2)."""
    d = X[elements[:, 1]] - X[elements[:, 0]]
    ln = np.hypot(d[:, 0], d[:, 1])
    normals = np.stack([-d[:, 1] / ln, d[:, 0] / ln], 1)
    tangents = np.stack([d[:, 0] / ln, d[:, 1] / ln], 1)
    return normals, tangents


def genghostnodes(nodes, normals, offset: float):
    """Ghost nodes at `offset` along the normals (src/genghostnodes.jl:1-6)."""
    return nodes + normals * offset


def processmesh(meshname: str, markernames):
    """Import a 2-D CGNS mesh (src/processmesh.jl:1-218).

    Returns (Y, vertices, Y_idx_in, Y_idx_bc, Y_idx_bc_g, cells, bc_normals,
    bc_tangents):
      Y          nodes: triangle centroids, boundary midpoints (CGNS element
                 order), then one ghost node per boundary node;
      vertices   the mesh vertices (P, 2);
      Y_idx_in   indices of the triangle centroids;
      Y_idx_bc   one index array per marker (its boundary midpoints);
      Y_idx_bc_g one index array per marker (its ghost nodes, same order);
      cells      (triangles (T, 3), boundary lines (B, 2)) vertex indices;
      bc_normals, bc_tangents  per marker, (k, 2) unit vectors, outward.
    """
    fid = H5File(meshname)
    dom = "Base/dom-1"
    verts = extractcoordinates(fid, f"{dom}/GridCoordinates/CoordinateX", f"{dom}/GridCoordinates/CoordinateY")
    tri_faces, tri_elements = extractelements(fid, verts, f"{dom}/TriElements/ElementConnectivity/ data", 3)
    int_range = fid.read(f"{dom}/TriElements/ElementRange/ data").astype(np.int64).ravel()
    Y_idx_in = np.arange(int_range[0] - 1, int_range[1])
    bc_range = [fid.read(f"{dom}/{m}/ElementRange/ data").astype(np.int64).ravel() for m in markernames]
    bc_max = max(int(r.max()) for r in bc_range)
    bc_min = min(int(r.min()) for r in bc_range)
    Y = np.full((bc_max, 2), np.nan)
    Y[Y_idx_in] = tri_faces
    Y_idx_bc, bc_lines, bc_normals, bc_tangents, bc_elems = [], [], [], [], []
    for m, r in zip(markernames, bc_range):
        idx = np.arange(r[0] - 1, r[1])
        lines, elems = extractelements(fid, verts, f"{dom}/{m}/ElementConnectivity/ data", 2)
        nrm, tng = calculatenormal(verts, elems)
        Y[idx] = lines
        Y_idx_bc.append(idx)
        bc_lines.append(lines)
        bc_normals.append(nrm)
        bc_tangents.append(tng)
        bc_elems.append(elems)
    # orient every normal outward: the first node of the first marker and its
    # nearest interior node fix the sign (src/processmesh.jl:125-139)
    tree_in = cKDTree(Y[Y_idx_in])
    b0 = Y[Y_idx_bc[0][0]]
    _, j = tree_in.query(b0, k=1)
    outward = b0 - Y[Y_idx_in][j]
    orient = np.sign(np.dot(outward, bc_normals[0][0]))
    bc_normals = [orient * v for v in bc_normals]
    bc_tangents = [orient * v for v in bc_tangents]
    # ghost offset: the mean distance from a boundary node to its nearest
    # interior node (src/processmesh.jl:141-168)
    dists = np.concatenate([tree_in.query(Y[idx], k=1)[0] for idx in Y_idx_bc])
    offset = abs(float(np.mean(dists)))
    shift = bc_max - (int_range[1])        # Y_idx_bc .+ Y_idx_bc_max .- int_range[2]
    Y_idx_bc_g = [idx + shift for idx in Y_idx_bc]
    ghosts = np.full((bc_max - bc_min + 1, 2), np.nan)
    for idx, lines, nrm in zip(Y_idx_bc, bc_lines, bc_normals):
        ghosts[idx - int_range[1]] = genghostnodes(lines, nrm, offset)
    Y = np.concatenate([Y, ghosts])
    cells = (tri_elements, np.concatenate(bc_elems) if bc_elems else np.zeros((0, 2), np.int64))
    return Y, verts, Y_idx_in, Y_idx_bc, Y_idx_bc_g, cells, bc_normals, bc_tangents
