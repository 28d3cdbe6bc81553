"""Independent FP64 NumPy twin of the reference plain-DC path, for tiny grids only.

TEST INFRASTRUCTURE. Only tests/ may import this module. It is written separately from
dc_oracle.cpp and follows the reference's data structures literally (per-cell maps of hermite
samples, process_edge per grid point, an SVD of A^T A for the solve), so that a slip in the C++
restatement shows up as a mismatch. Pure-Python loops: keep grids at or below ~14^3.

Reference (file:line in the reference tree):
  src/cpp/dc_backend.cpp:51-125      CpuDualContouringBackend::extract
  src/cpp/contouring.cpp:141-271     gradientAt, generate_cells
  src/cpp/contouring.cpp:486-741     extract_mesh_from_cells (hermite overload)
  src/cpp/Cell.cpp:49-361, 586-600   lerp_t, has_vertex, trilinear gradient, hermite data, solve
  src/cpp/hermite_update.cpp:34-438  corner offsets, edge ids, process_edge, averaging
The solve uses numpy.linalg.svd of A^T A and forms V diag(1/s) U^T exactly as Cell.cpp:335-358
does with Eigen's JacobiSVD, independently of the oracle's Jacobi eigen-solver.
"""
import math

import numpy as np

EDGE_PAIRS = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
              (0, 4), (1, 5), (2, 6), (3, 7)]
CORNER_OFF = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1),
              (0, 1, 1)]
EDGE_FACES = [(3, 5), (0, 5), (2, 5), (1, 5), (3, 4), (0, 4), (2, 4), (1, 4), (1, 3), (0, 3),
              (0, 2), (1, 2)]
DIRS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def _f32(x):
    return np.float32(x)


def grid_coord(o, i, v):
    """dc_backend.cpp:63-65: FP32 multiply, then FP32 add."""
    return float(_f32(_f32(o) + _f32(_f32(i) * _f32(v))))


def _sq(v):
    return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]


def _normalized(v):
    z = _sq(v)
    if z > 0.0:
        s = math.sqrt(z)
        return [v[0] / s, v[1] / s, v[2] / s]
    return list(v)


class Cell:
    def __init__(self, i, j, k, corners, sdf):
        self.ix, self.iy, self.iz = i, j, k
        self.corners = corners
        self.sdf = sdf
        self.has_vertex = any(s < 0 for s in sdf) and any(s > 0 for s in sdf)
        self.hp = {}
        self.hn = {}
        self.vertex = [0.0, 0.0, 0.0]

    def grad_trilinear(self, p):
        S000, S100, S110, S010, S001, S101, S111, S011 = self.sdf
        x0, x1 = self.corners[0][0], self.corners[1][0]
        y0, y1 = self.corners[0][1], self.corners[3][1]
        z0, z1 = self.corners[0][2], self.corners[4][2]
        dx, dy, dz = x1 - x0, y1 - y0, z1 - z0
        u = v = w = 0.5
        if abs(dx) > 1e-12:
            u = (p[0] - x0) / dx
        if abs(dy) > 1e-12:
            v = (p[1] - y0) / dy
        if abs(dz) > 1e-12:
            w = (p[2] - z0) / dz
        u, v, w = (max(0.0, min(1.0, t)) for t in (u, v, w))
        omu, omv, omw = 1.0 - u, 1.0 - v, 1.0 - w
        du = ((S100 - S000) * omv * omw + (S110 - S010) * v * omw + (S101 - S001) * omv * w
              + (S111 - S011) * v * w)
        dv = ((S010 - S000) * omu * omw + (S110 - S100) * u * omw + (S011 - S001) * omu * w
              + (S111 - S101) * u * w)
        dw = ((S001 - S000) * omu * omv + (S101 - S100) * u * omv + (S011 - S010) * omu * v
              + (S111 - S110) * u * v)
        g = [0.0, 0.0, 0.0]
        if abs(dx) > 1e-12:
            g[0] = du / dx
        if abs(dy) > 1e-12:
            g[1] = dv / dy
        if abs(dz) > 1e-12:
            g[2] = dw / dz
        return g

    def fill_hermite(self):
        if not self.has_vertex:
            return
        for e, (a, b) in enumerate(EDGE_PAIRS):
            fa, fb = self.sdf[a], self.sdf[b]
            if fa * fb > 0.0:
                continue
            t = fa / (fa - fb + 1e-16)
            pa, pb = self.corners[a], self.corners[b]
            p = [(1.0 - t) * pa[d] + t * pb[d] for d in range(3)]
            n = self.grad_trilinear(p)
            if math.sqrt(_sq(n)) > 0.0:
                n = _normalized(n)
            self.hp[e] = p
            self.hn[e] = n

    def update(self):
        if not self.has_vertex or not self.hp:
            self.has_vertex = False
            return
        rows, rhs = [], []
        for e in sorted(self.hp):
            p = self.hp[e]
            n = list(self.hn[e])
            if math.sqrt(_sq(n)) > 0:
                n = _normalized(n)
            rows.append(n)
            rhs.append((n[0] * p[0] + n[1] * p[1]) + n[2] * p[2])
        A = np.array(rows, dtype=np.float64)
        b = np.array(rhs, dtype=np.float64)
        ATA = A.T @ A
        ATb = A.T @ b
        U, S, Vt = np.linalg.svd(ATA)
        Sinv = np.array([1.0 / s if s > 0.01 else 0.0 for s in S])
        pinv = Vt.T @ np.diag(Sinv) @ U.T
        centroid = np.mean(np.array([self.hp[e] for e in sorted(self.hp)]), axis=0)
        c = pinv @ (ATb - ATA @ centroid)
        self.vertex = list(centroid + c)


def corner_offsets(c):
    dz = 1 if c >= 4 else 0
    rem = c % 4
    return (1 if rem in (1, 2) else 0), (1 if rem in (2, 3) else 0), dz


def edge_idx(cell, p0, p1):
    d0 = (p0[0] - cell.ix, p0[1] - cell.iy, p0[2] - cell.iz)
    d1 = (p1[0] - cell.ix, p1[1] - cell.iy, p1[2] - cell.iz)
    for e, (a, b) in enumerate(EDGE_PAIRS):
        A, B = corner_offsets(a), corner_offsets(b)
        if (d0 == A and d1 == B) or (d0 == B and d1 == A):
            return e
    return -1


def gradient_at(GV, S, i, j, k, res):
    nx, ny, nz = res
    cl = lambda v, hi: max(0, min(hi, v))
    im, ip = cl(i - 1, nx - 1), cl(i + 1, nx - 1)
    jm, jp = cl(j - 1, ny - 1), cl(j + 1, ny - 1)
    km, kp = cl(k - 1, nz - 1), cl(k + 1, nz - 1)
    g = [0.0, 0.0, 0.0]
    dx = GV[ip][j][k][0] - GV[im][j][k][0]
    if abs(dx) > 1e-12:
        g[0] = (S[ip][j][k] - S[im][j][k]) / dx
    dy = GV[i][jp][k][1] - GV[i][jm][k][1]
    if abs(dy) > 1e-12:
        g[1] = (S[i][jp][k] - S[i][jm][k]) / dy
    dz = GV[i][j][kp][2] - GV[i][j][km][2]
    if abs(dz) > 1e-12:
        g[2] = (S[i][j][kp] - S[i][j][km]) / dz
    nrm = math.sqrt(_sq(g))
    if nrm > 0.0:
        g = [g[0] / nrm, g[1] / nrm, g[2] / nrm]
    return g


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def extract(values, nx, ny, nz, vx, vy, vz, ox, oy, oz):
    """Returns (V float64 (A,3), F int32 (F,4), active cell linear ids)."""
    vals = np.asarray(values, dtype=np.float64).reshape(nz, ny, nx)
    S = [[[float(vals[k, j, i]) for k in range(nz)] for j in range(ny)] for i in range(nx)]
    GV = [[[(grid_coord(ox, i, vx), grid_coord(oy, j, vy), grid_coord(oz, k, vz))
            for k in range(nz)] for j in range(ny)] for i in range(nx)]
    cx, cy, cz = nx - 1, ny - 1, nz - 1
    cidx = lambda i, j, k: i + cx * (j + cy * k)
    cells = [None] * (cx * cy * cz)
    for k in range(cz):
        for j in range(cy):
            for i in range(cx):
                corners, sdf = [], []
                for (a, b, c) in CORNER_OFF:
                    corners.append(GV[i + a][j + b][k + c])
                    sdf.append(S[i + a][j + b][k + c])
                cell = Cell(i, j, k, corners, sdf)
                cell.fill_hermite()
                cells[cidx(i, j, k)] = cell

    inside = lambda a, b, c: 0 < a < nx - 1 and 0 < b < ny - 1 and 0 < c < nz - 1
    records = []

    def process_edge(p0, p1):
        if not inside(*p0) and not inside(*p1):
            return
        idx = cidx(*p0)
        if idx < 0 or idx >= len(cells):
            return
        cell = cells[idx]
        e = edge_idx(cell, p0, p1)
        if e == -1 or e not in cell.hp:
            return
        d1, d2 = DIRS[EDGE_FACES[e][0]], DIRS[EDGE_FACES[e][1]]
        n1 = (cell.ix + d1[0], cell.iy + d1[1], cell.iz + d1[2])
        n2 = (cell.ix + d2[0], cell.iy + d2[1], cell.iz + d2[2])
        n3 = (n1[0] + d2[0], n1[1] + d2[1], n1[2] + d2[2])
        group = [idx, cidx(*n1), cidx(*n2), cidx(*n3)]
        if not all(cells[q].has_vertex for q in group):
            return
        records.append((group, [edge_idx(cells[q], p0, p1) for q in group]))

    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if i < nx - 1:
                    process_edge((i, j, k), (i + 1, j, k))
                if j < ny - 1:
                    process_edge((i, j, k), (i, j + 1, k))
                if k < nz - 1:
                    process_edge((i, j, k), (i, j, k + 1))
    for group, eids in records:
        acc = [0.0, 0.0, 0.0]
        for q, e in zip(group, eids):
            n = cells[q].hn[e]
            if math.sqrt(_sq(n)) > 0:
                acc = [acc[0] + n[0], acc[1] + n[1], acc[2] + n[2]]
        acc = _normalized(acc)
        for q, e in zip(group, eids):
            cells[q].hn[e] = list(acc)

    for c in cells:
        c.update()

    vid = [-1] * len(cells)
    V, active = [], []
    for k in range(cz):
        for j in range(cy):
            for i in range(cx):
                c = cells[cidx(i, j, k)]
                if c.has_vertex:
                    vid[cidx(i, j, k)] = len(V)
                    V.append(c.vertex)
                    active.append(cidx(i, j, k))

    def cv(i, j, k):
        if 0 <= i < cx and 0 <= j < cy and 0 <= k < cz:
            return vid[cidx(i, j, k)]
        return -1

    F = []

    def quad(c, g0p, g1p):
        c0, c1, c2, c3 = c
        if min(c) < 0:
            return
        g0 = gradient_at(GV, S, *g0p, (nx, ny, nz))
        g1 = gradient_at(GV, S, *g1p, (nx, ny, nz))
        gn = [g0[0] + g1[0], g0[1] + g1[1], g0[2] + g1[2]]
        pa, pb, pc, pd = V[c0], V[c1], V[c2], V[c3]
        qn = _cross([pb[d] - pa[d] for d in range(3)], [pd[d] - pa[d] for d in range(3)])
        if _sq(qn) < 1e-24:
            qn = _cross([pc[d] - pa[d] for d in range(3)], [pd[d] - pa[d] for d in range(3)])
        if _sq(gn) > 0 and _sq(qn) > 0:
            if (qn[0] * gn[0] + qn[1] * gn[1]) + qn[2] * gn[2] < 0.0:
                c1, c3 = c3, c1
        F.append([c0, c1, c2, c3])

    for i in range(cx):
        for j in range(1, cy):
            for k in range(1, cz):
                if S[i][j][k] * S[i + 1][j][k] < 0.0:
                    quad((cv(i, j - 1, k - 1), cv(i, j, k - 1), cv(i, j, k), cv(i, j - 1, k)),
                         (i, j, k), (i + 1, j, k))
    for i in range(1, cx):
        for j in range(cy):
            for k in range(1, cz):
                if S[i][j][k] * S[i][j + 1][k] < 0.0:
                    quad((cv(i - 1, j, k - 1), cv(i, j, k - 1), cv(i, j, k), cv(i - 1, j, k)),
                         (i, j, k), (i, j + 1, k))
    for i in range(1, cx):
        for j in range(1, cy):
            for k in range(cz):
                if S[i][j][k] * S[i][j][k + 1] < 0.0:
                    quad((cv(i - 1, j - 1, k), cv(i, j - 1, k), cv(i, j, k), cv(i - 1, j, k)),
                         (i, j, k), (i, j, k + 1))
    return (np.array(V, dtype=np.float64).reshape(-1, 3), np.array(F, dtype=np.int32).reshape(-1, 4),
            np.array(active, dtype=np.int64))
