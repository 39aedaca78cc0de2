"""Structured quad4/hex8 meshes in the reference's node numbering.

A reference mesh file (``Ascii2D``/``Ascii3D``, include/io/io_mesh_ascii.tcc:4-143)
lists node coordinates and, per cell, its node ids in the element's local
order: quadrilateral CCW from the (x-, y-) corner
(include/elements/2d/quadrilateral_element.tcc:1-7), hexahedron bottom face CCW
then top face (include/elements/3d/hexahedron_element.tcc:1-13). These helpers
produce the same arrays for an axis-aligned grid with nodes numbered x-fastest,
then y, then z, and cells likewise.
"""
import numpy as np


def structured_mesh(origin, spacing, ncells):
    """Return ``(coords[nnodes, dim], cells[ncells, 2**dim])``."""
    origin = np.asarray(origin, dtype=np.float64)
    spacing = np.asarray(spacing, dtype=np.float64)
    n = [int(v) for v in ncells]
    dim = len(n)
    if dim not in (2, 3) or origin.shape != (dim,) or spacing.shape != (dim,):
        raise ValueError("structured_mesh needs matching 2D or 3D origin/spacing/ncells")
    nn = [v + 1 for v in n]
    axes = [origin[k] + np.arange(nn[k], dtype=np.float64) * spacing[k] for k in range(dim)]
    if dim == 2:
        gy, gx = np.meshgrid(axes[1], axes[0], indexing="ij")
        coords = np.stack([gx.ravel(), gy.ravel()], axis=1)
        j, i = np.meshgrid(np.arange(n[1]), np.arange(n[0]), indexing="ij")
        i = i.ravel().astype(np.int64)
        j = j.ravel().astype(np.int64)
        nid = lambda a, b: a + nn[0] * b
        cells = np.stack([nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)], axis=1)
    else:
        gz, gy, gx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        coords = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
        k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
        i = i.ravel().astype(np.int64)
        j = j.ravel().astype(np.int64)
        k = k.ravel().astype(np.int64)
        nid = lambda a, b, c: a + nn[0] * (b + nn[1] * c)
        cells = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                          nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                          nid(i, j + 1, k + 1)], axis=1)
    return np.ascontiguousarray(coords), np.ascontiguousarray(cells)


def cells_in_box(coords, cells, lo, hi):
    """Ids of cells whose centroid lies in the axis-aligned box [lo, hi]."""
    c = coords[cells].mean(axis=1)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    inside = np.all((c >= lo) & (c <= hi), axis=1)
    return np.nonzero(inside)[0].astype(np.int64)


def nodes_where(coords, axis, value, tol=1e-12):
    """Ids of nodes with ``coords[:, axis] == value`` (within ``tol``)."""
    return np.nonzero(np.abs(coords[:, axis] - value) <= tol)[0].astype(np.int64)
