"""Lifting the remeshed sites back to 3D: point location in the parameterized
mesh on the GPU (cvt_locate_points, include/cvt.h) and the reference's
recover() step on arrays.

The reference's recover() (source/recover.h:83-153) takes the 2D mesh, its
3D counterpart, the CDT points (sites first, constraint vertices last) and
the CDT triangles, and builds the output surface mesh:
  * each site is located in the 2D mesh (the first face whose barycentric
    weights are all >= 0, tested face by face) and lifted with those
    weights;
  * each constraint vertex takes its own 3D position;
  * each CDT triangle is kept when its centroid lies in some face.
`recover` here does the same with arrays; the face search runs on the GPU.
"""
from __future__ import annotations

import numpy as np

from ._lib import call
from .engine import _out_ptr, _ptr


def locate_points(uv, tri, q, p3=None, out=None):
    """First face (in face order) holding each query point, -1 for none.

    uv: (nv, 2) float64 vertex coordinates; tri: (nf, 3) int32 vertex
    indices; q: (nq, 2) float64 points; p3: optional (nv, 3) float64 3D
    vertices.  Returns (face (nq,) int32, weights (nq, 3) float64) and, with
    p3, the interpolated 3D points (nq, 3).  numpy arrays or torch tensors
    (host or device memory) are accepted.  out: None for new numpy arrays,
    or the arrays/tensors to write (host or device memory), (face, w) or,
    with p3, (face, w, xyz)."""
    uvp, k0 = _ptr(uv, np.float64)
    trp, k1 = _ptr(tri, np.int32)
    qp, k2 = _ptr(q, np.float64)
    p3p, k3 = _ptr(p3, np.float64)
    nv = int(np.prod(k0.shape)) // 2 if k0 is not None else 0
    nf = int(np.prod(k1.shape)) // 3 if k1 is not None else 0
    nq = int(np.prod(k2.shape)) // 2
    if out is None:
        face = np.empty(nq, np.int32)
        w = np.empty((nq, 3), np.float64)
        out3 = np.empty((nq, 3), np.float64) if p3 is not None else None
    else:
        if len(out) != (2 if p3 is None else 3):
            raise ValueError("out must be (face, w) or, with p3, (face, w, xyz)")
        face, w = out[0], out[1]
        out3 = out[2] if p3 is not None else None
    pf = _out_ptr(face, np.int32, nq)
    pw = _out_ptr(w, np.float64, 3 * nq)
    po = None if out3 is None else _out_ptr(out3, np.float64, 3 * nq)
    call("cvt_locate_points", uvp, p3p, nv, trp, nf, qp, nq, pf, pw, po)
    return (face, w) if p3 is None else (face, w, out3)


def centroids(points, triangles):
    """Triangle centroids ((a + b + c) / 3 per axis, CGAL::centroid's
    formula for three points)."""
    p = np.asarray(points, np.float64).reshape(-1, 2)
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    return (p[t[:, 0]] + p[t[:, 1]] + p[t[:, 2]]) / 3.0


def recover(uv, p3, tri, points, triangles, num_constraints, constraint_vertices):
    """The reference's recover() on arrays (source/recover.h:83-153).

    uv, p3, tri: the parameterized mesh (2D vertices, their 3D positions,
    faces); points: (n, 2) CDT points whose last num_constraints entries are
    constraint vertices; triangles: (m, 3) CDT triangles indexing points;
    constraint_vertices: the mesh vertex of each constraint point.  Returns
    (vertices (n, 3), faces (k, 3) int32): the lifted sites, then the
    constraint vertices' own 3D positions, and the CDT triangles whose
    centroid lies in the mesh.  A site outside every face takes the location
    of the site before it, as in the reference (its locate() result is not
    checked at :111); a first site outside every face lifts to 0."""
    pts = np.asarray(points, np.float64).reshape(-1, 2)
    ns = pts.shape[0] - int(num_constraints)
    face, w, lifted = locate_points(uv, tri, pts[:ns], p3=p3)
    tri_a = np.asarray(tri, np.int64).reshape(-1, 3)
    p3a = np.asarray(p3, np.float64).reshape(-1, 3)
    verts = np.empty((pts.shape[0], 3), np.float64)
    verts[:ns] = lifted
    for k in np.flatnonzero(face < 0):
        # the stale location: the nearest earlier site that was located
        j = k - 1
        while j >= 0 and face[j] < 0:
            j -= 1
        if j < 0:
            verts[k] = 0.0
            continue
        t = tri_a[face[j]]
        r = np.zeros(3)
        for jj in range(3):
            r += p3a[t[jj]] * w[j, jj]
        verts[k] = r
    verts[ns:] = p3a[np.asarray(constraint_vertices, np.int64)]
    tr = np.asarray(triangles, np.int32).reshape(-1, 3)
    cf, _ = locate_points(uv, tri, centroids(pts, tr))
    return verts, tr[cf >= 0]
