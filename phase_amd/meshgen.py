"""Synthetic mesh and coefficient generators for the pressure-solve workloads.

Meshes are returned in the form FiniteVolumeGrid2D::init(nodes, cptr, cind)
takes (src/2D/Unstructured/FiniteVolumeGrid2D/FiniteVolumeGrid2D.cpp:20-35):
node coordinates (n_nodes x 2, float64), and a cell -> node CSR (cptr, cind,
int32) listing each cell's nodes counter-clockwise.

* rectilinear(): StructuredRectilinearGrid::init
  (StructuredRectilinearGrid.cpp:40-95): nodes j-major, cells j-major, each
  cell (i,j) -> nodes (i,j), (i+1,j), (i+1,j+1), (i,j+1).
* delaunay_square(): a Delaunay triangulation of the unit square from k
  equally spaced points per edge plus uniform interior points (Qhull through
  scipy.spatial, triangles oriented counter-clockwise).  With h = 4k hull
  points and n points in all, the mesh has 2n - 2 - h triangles.
* lognormal_face_coeffs(), zero_mean_normal(): seeded numpy PCG64 draws.
"""
import numpy as np


def rectilinear(nx, ny, width=1.0, height=1.0):
    hx0, hy0 = width / nx, height / ny
    xd = np.array([i * hx0 for i in range(nx + 1)])
    yd = np.array([j * hy0 for j in range(ny + 1)])
    X, Y = np.meshgrid(xd, yd)  # j-major
    xy = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i, j = i.ravel(), j.ravel()
    nnx = nx + 1
    cind = np.stack([j * nnx + i, j * nnx + i + 1, (j + 1) * nnx + i + 1, (j + 1) * nnx + i],
                    axis=1).ravel().astype(np.int32)
    cptr = (4 * np.arange(nx * ny + 1)).astype(np.int32)
    return xy, cptr, cind


def delaunay_square(k, n_interior, seed=1):
    """Delaunay mesh of [0,1]^2: 4k boundary points (k per edge) + n_interior
    uniform points drawn from PCG64(seed)."""
    from scipy.spatial import Delaunay

    t = np.arange(k) / k
    z, o = np.zeros(k), np.ones(k)
    bnd = np.concatenate([np.stack([t, z], 1), np.stack([o, t], 1),
                          np.stack([1.0 - t, o], 1), np.stack([z, 1.0 - t], 1)])
    rng = np.random.Generator(np.random.PCG64(seed))
    inner = rng.random((n_interior, 2))
    # keep interior points strictly inside (a draw of exactly 0 is possible)
    inner = np.clip(inner, 1e-12, 1.0 - 1e-12)
    xy = np.concatenate([bnd, inner]).astype(np.float64)
    tri = Delaunay(xy).simplices.astype(np.int64)
    p0, p1, p2 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    cross = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0])
    flip = cross < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    keep = np.abs(cross) > 0
    tri = tri[keep]
    cind = tri.astype(np.int32).ravel()
    cptr = (3 * np.arange(len(tri) + 1)).astype(np.int32)
    return xy, cptr, cind


def delaunay_cells_for(n_cells_target, k):
    """Interior point count giving exactly n_cells_target triangles with k
    points per edge (T = 2n - 2 - h, h = 4k)."""
    h = 4 * k
    n = (n_cells_target + 2 + h) // 2
    return n - h


def boundary_patch(face_lnode, face_rnode, face_rcell, xy, x0=0.0, x1=1.0, y0=0.0, y1=1.0):
    """Patch id per face of an axis-aligned box: -1 interior, 0 'x-', 1 'x+',
    2 'y-', 3 'y+' (the StructuredRectilinearGrid default patches,
    StructuredRectilinearGrid.cpp:175-194)."""
    a, b = xy[face_lnode], xy[face_rnode]
    pid = np.full(len(face_lnode), -1, np.int32)
    bnd = face_rcell < 0
    pid[bnd & (a[:, 0] == x0) & (b[:, 0] == x0)] = 0
    pid[bnd & (a[:, 0] == x1) & (b[:, 0] == x1)] = 1
    pid[bnd & (a[:, 1] == y0) & (b[:, 1] == y0)] = 2
    pid[bnd & (a[:, 1] == y1) & (b[:, 1] == y1)] = 3
    return pid


def lognormal_face_coeffs(n_faces, sigma, seed=2):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(sigma * rng.standard_normal(n_faces))


def zero_mean_normal(n, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    b = rng.standard_normal(n)
    return b - b.mean()
