"""Per-vertex density weights of the parameterized mesh on the GPU
(cvt_vertex_weights, include/cvt.h): the vertex weights the rasterizer
interpolates.

The reference's weighting_vertices_with_area (source/weighting.h:37-97)
gives each 2D vertex the ratio of the 3D area to the 2D area of its faces
(both summed in face order), and replaces weights above 1e10 (vertices
whose faces are flat in the plane) by the largest other weight.
"""
from __future__ import annotations

import numpy as np

from ._lib import call
from .engine import _out_ptr, _ptr


def vertex_weights(uv, p3, tri, out=None):
    """Density weight of every vertex.

    uv: (nv, 2) float64 2D vertices; p3: (nv, 3) float64 3D position of each
    2D vertex; tri: (nf, 3) int32 faces in vertices_around_face order.
    numpy arrays or torch tensors (host or device memory) are accepted.
    Returns (nv,) float64: area3 / area2 summed over each vertex's faces,
    with the > 1e10 rule; NaN for a vertex in no face.  out: an (nv,)
    float64 array or tensor to write (host or device; a device tensor keeps
    the weights on the GPU for the rasterizer), or None for a new numpy
    array."""
    uvp, k0 = _ptr(uv, np.float64)
    p3p, k1 = _ptr(p3, np.float64)
    trp, k2 = _ptr(tri, np.int32)
    nv = int(np.prod(k0.shape)) // 2
    if int(np.prod(k1.shape)) != 3 * nv:
        raise ValueError("p3 must hold one 3D point per uv vertex")
    nf = int(np.prod(k2.shape)) // 3
    if out is None:
        out = np.empty(nv, np.float64)
    call("cvt_vertex_weights", uvp, p3p, nv, trp, nf, _out_ptr(out, np.float64, nv))
    return out
