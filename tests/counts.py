"""Mesh counts of plain dual contouring computed directly from the samples with NumPy (an
independent check of the marking pass and the scans at sizes where the oracle is slow).

- active cells: cells with a corner < 0 and a corner > 0 (Cell.cpp:57-77);
- quads: interior grid edges with a strict sign change (contouring.cpp:607-729). Such an edge's
  endpoints are a negative and a positive corner of each of its 4 incident cells, so all 4 are
  active and the edge always emits its quad.

Works on z-slabs so that a 1024^3 grid needs a few GB beyond its samples.
"""
import numpy as np


def mesh_counts(values, nx, ny, nz, slab=64):
    v = np.asarray(values, dtype=np.float32).reshape(nz, ny, nx)
    active = 0
    faces = 0
    for k0 in range(0, nz - 1, slab):
        k1 = min(nz - 1, k0 + slab)          # cell layers [k0, k1)
        s = v[k0:k1 + 1]                     # sample layers [k0, k1]
        neg, pos = s < 0, s > 0
        # cells: any corner over z, then y, then x pairs
        nz_ = neg[:-1] | neg[1:]
        pz_ = pos[:-1] | pos[1:]
        ny_ = nz_[:, :-1] | nz_[:, 1:]
        py_ = pz_[:, :-1] | pz_[:, 1:]
        nx_ = ny_[:, :, :-1] | ny_[:, :, 1:]
        px_ = py_[:, :, :-1] | py_[:, :, 1:]
        active += int(np.count_nonzero(nx_ & px_))
        del nz_, pz_, ny_, py_, nx_, px_
        # quads on edges whose lower endpoint has k in [k0, k1) (each edge counted once)
        lo_n, lo_p = neg[:-1], pos[:-1]
        # x edges: i in [0, cx), j in [1, cy), k in [1, cz)
        kx = slice(max(1, k0) - k0, k1 - k0)
        ex = (lo_n[:, :, :-1] & lo_p[:, :, 1:]) | (lo_p[:, :, :-1] & lo_n[:, :, 1:])
        faces += int(np.count_nonzero(ex[kx, 1:ny - 1, :]))
        # y edges: i in [1, cx), j in [0, cy), k in [1, cz)
        ey = (lo_n[:, :-1, :] & lo_p[:, 1:, :]) | (lo_p[:, :-1, :] & lo_n[:, 1:, :])
        faces += int(np.count_nonzero(ey[kx, :, 1:nx - 1]))
        # z edges: i in [1, cx), j in [1, cy), k in [0, cz)
        ez = (neg[:-1] & pos[1:]) | (pos[:-1] & neg[1:])
        faces += int(np.count_nonzero(ez[:, 1:ny - 1, 1:nx - 1]))
    return active, faces
