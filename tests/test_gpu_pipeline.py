"""The stages of the reference's remeshing run after the parameterization
(source/main.cpp:196-241), chained on the GPU through the C ABI, each
compared with the oracle's chain on the same inputs:

  vertex weights (weighting.h:37-97) -> density raster (discretization.h:
  69-128) -> constraint mask (gcvt.h:143-159) and reference site
  initialisation (gcvt.h:76-122, host) -> gCVT (gcvt.cu:1087-1156) -> CDT
  input points (delaunay.h:48-58) -> the constrained Delaunay triangulation
  of the sites and the patch boundary (delaunay.h:60-79, GpuDel::compute at
  main.cpp:236) -> the site lift and the triangle filter of recover()
  (recover.h:63-141).

Bar: every stage before the triangulation bit-exact; the triangulation
checked as a constrained Delaunay triangulation (every boundary edge
present, hull tiled, empty circles); the next stage takes the GPU's
output."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,num,m,seed", [(128, 120, 20, 11), (256, 500, 40, 12)])
def test_pipeline_mesh_to_lifted_sites(gpu, orc, n, num, m, seed):
    from cvtamd import compat, recover, weighting, workloads as wl
    from test_gpu_compat import reference_init
    uv, p3, tri, _ = wl.patch_mesh(seed=seed, m=m)
    # weights of the 2D vertices
    w = weighting.vertex_weights(uv, p3, tri)
    assert np.array_equal(w, orc.vertex_weights(uv, p3, tri))
    assert np.all(w >= 1.0 - 1e-12)                  # a graph surface never shrinks area
    # discretization(): bounds, scale and shift (discretization.h:89-105)
    l, b = uv[:, 0].min(), uv[:, 1].min()
    ext = max(uv[:, 0].max() - l, uv[:, 1].max() - b)
    scale = ext / (n - 1)
    pts = np.ascontiguousarray(uv - [l, b])
    dens = compat.rasterize(pts, w, tri, scale, n)
    assert np.array_equal(dens, orc.raster(pts, w, tri, scale, n))
    # generateMask: constraint vertices (here the patch corners) to pixels
    mask = np.zeros((n, n), np.uint8)
    for v in (0, m - 1, m * (m - 1), m * m - 1):
        x, y = int((uv[v, 0] - l) / scale), int((uv[v, 1] - b) / scale)
        mask[y, x] = 1
    vor = reference_init(orc, dens, mask, num)
    got, ref = vor.copy(), vor.copy()
    it, ns = compat.gcvt(got, dens, mask, 1, 300)
    it_r, _, ns_r = orc.compat_lloyd(dens, ref, mask, max_iter=300)
    assert (it, ns) == (it_r, ns_r) and np.array_equal(got, ref)
    # CDT input points: the sites, in the mesh's coordinates
    sites = compat.extract_points(got, mask, scale, l, b)
    assert np.array_equal(sites, orc.extract_points(ref, mask, scale, l, b))
    assert len(sites) == ns - int(mask.sum())
    # recover(): every site lies in the patch and is lifted by its face
    face, wts, xyz = recover.locate_points(uv, tri, sites, p3=p3)
    f_r, w_r, x_r = orc.locate(uv, tri, sites, p3=p3)
    assert np.array_equal(face, f_r) and np.array_equal(wts, w_r) and np.array_equal(xyz, x_r)
    assert np.all(face >= 0)
    assert np.abs(xyz[:, 2] - 0.3 * np.sin(3 * xyz[:, 0]) * np.cos(2 * xyz[:, 1])).max() < 0.05
    # the CDT input: the sites, then the boundary vertices as constraint
    # points joined by the boundary edges (delaunayInput, delaunay.h:60-79)
    from cvtamd.delaunay import delaunay
    from test_gpu_delaunay import check_constrained
    ring = ([i for i in range(m)] + [j * m + m - 1 for j in range(1, m)] +
            [(m - 1) * m + i for i in range(m - 2, -1, -1)] + [j * m for j in range(m - 2, 0, -1)])
    bpts = uv[ring]
    # a site exactly on a boundary vertex would be the same point twice
    dup = (sites[:, None, :] == bpts[None, :, :]).all(-1).any(1)
    cs = sites[~dup]
    cpts = np.concatenate([cs, bpts])
    segs = np.stack([len(cs) + np.arange(len(ring)), len(cs) + (np.arange(len(ring)) + 1) % len(ring)],
                    1).astype(np.int32)
    ctri, copp = delaunay(cpts, segments=segs)
    check_constrained(cpts, segs, ctri, copp)
    # recover() keeps the triangles whose centroid lies in the mesh
    # (recover.h:134-141).  The patch is convex, so every triangle of positive
    # area keeps its centroid inside it; the zero-area ones along the
    # collinear boundary have theirs on the boundary, where the barycentric
    # test may round either way (recover() would drop those)
    cen = cpts[ctri].mean(axis=1)
    cface = recover.locate_points(uv, tri, cen)[0]
    a, b_, c = cpts[ctri[:, 0]], cpts[ctri[:, 1]], cpts[ctri[:, 2]]
    area2 = (b_[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b_[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    inside = (cen[:, 0] > 0) & (cen[:, 0] < 1) & (cen[:, 1] > 0) & (cen[:, 1] < 1)
    assert np.all(cface[inside & (area2 > 0)] >= 0)
    assert (area2 > 0).sum() >= len(ctri) - 4 * (m - 1)
