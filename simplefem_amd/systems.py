"""Synthetic CSR systems of the kind FEEC++ hands to its CG solvers.

* ``fd_laplacian_2d``      the 5-point finite-difference Laplacian of
                           tests/solver/solvespd_csr.cpp:157-177 (h = 1/(N+1)).
* ``poisson2d_p1``         P1 Lagrange stiffness on the unit square of
                           mesh/examples2D.hpp:194-211 (two triangles sharing
                           the (0,0)-(1,1) diagonal) after ``levels`` uniform
                           refinements: a (2^L+1)^2 vertex grid.
* ``poisson3d_p1_kuhn``    P1 Lagrange stiffness on n^3 cubes, each split into
                           the six Kuhn tetrahedra of mesh/examples3D.hpp:14-37
                           (all sharing the (0,0,0)-(1,1,1) diagonal).

The FEM matrices keep the structural pattern of the assembly
(fem/lagrangematrices.cpp:210-271 before exact zeros are dropped): every
vertex pair that shares an element has an entry, entries that touch a
Dirichlet (boundary) vertex are 0, and so is the diagonal of a Dirichlet row,
which makes the SSOR-family solvers skip that row.  Element contributions are
summed in exact integer arithmetic (units of 1/2 in 2D, h/6 in 3D) and
rounded once, so couplings that vanish are exact zeros.

Vertices are numbered lexicographically (x fastest) by default; with
numbering="reference" they are numbered as the reference's uniform
refinement numbers them (``reference_permutation``, a restatement of
mesh/mesh.simplicial2D.cpp:2205-2750 and mesh/mesh.simplicial3D.ur.cxx), which
gives the reference's CSR structure bit for bit: the structural pattern of
LagrangeCoefficientStiffnessMatrix (fem/lagrangematrices.cpp:547-782) and, with
exact zeros dropped, MatrixCSR(LagrangeStiffnessMatrix) + compressentries
(fem/lagrangematrices.cpp:210-271, tests/sullivan2D/poissondirichletlshaped.cpp:128-131),
whose values these match too.

``rhs_uniform`` draws the seeded U[-1,1) right-hand side (SplitMix64 stream)
on the active rows and puts 0 on Dirichlet rows, as the reference's RHS
assembly does (tests/sullivan2D/poissondirichletlshaped.cpp:149-176).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class CSRSystem:
    name: str
    n: int
    A: np.ndarray          # int32 row offsets (n+1)
    C: np.ndarray          # int32 column indices
    V: np.ndarray          # float64 values
    dirichlet: np.ndarray  # bool per row

    @property
    def nnz(self) -> int:
        return int(self.A[-1])

    def diagonal(self) -> np.ndarray:
        rows = np.repeat(np.arange(self.n), np.diff(self.A))
        d = np.zeros(self.n)
        on = self.C == rows
        np.add.at(d, rows[on], self.V[on])
        return d

    def dense(self) -> np.ndarray:
        M = np.zeros((self.n, self.n))
        rows = np.repeat(np.arange(self.n), np.diff(self.A))
        np.add.at(M, (rows, self.C), self.V)
        return M


def _coo_to_csr(n, rows, cols, vals):
    """Sums duplicates, sorts columns inside each row."""
    import scipy.sparse as sp

    m = sp.coo_matrix((vals, (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64))),
                      shape=(n, n)).tocsr()
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data


def fd_laplacian_2d(N: int) -> CSRSystem:
    """tests/solver/solvespd_csr.cpp:157-177: (4, -1, -1, -1, -1)/h^2 on an N x N grid."""
    h = 1.0 / (N + 1)
    h2 = h * h
    rows, cols, vals = [], [], []
    for e in range(N * N):
        x, y = divmod(e, N)
        rows.append(e); cols.append(e); vals.append(4.0 / h2)
        if x != 0:
            rows.append(e); cols.append((x - 1) * N + y); vals.append(-1.0 / h2)
        if x != N - 1:
            rows.append(e); cols.append((x + 1) * N + y); vals.append(-1.0 / h2)
        if y != 0:
            rows.append(e); cols.append(x * N + y - 1); vals.append(-1.0 / h2)
        if y != N - 1:
            rows.append(e); cols.append(x * N + y + 1); vals.append(-1.0 / h2)
    A, C, V = _coo_to_csr(N * N, np.array(rows), np.array(cols), np.array(vals))
    return CSRSystem(f"fd2d_N{N}", N * N, A, C, V, np.zeros(N * N, dtype=bool))


def _assemble(n_vert, elems, kref_int, scale, dirichlet):
    """elems: (E, nv) vertex ids of each element of one reference type;
    kref_int: integer element matrix (nv, nv) in units of `scale`."""
    nv = elems.shape[1]
    rows = np.repeat(elems, nv, axis=1).ravel()
    cols = np.tile(elems, (1, nv)).ravel()
    vals = np.tile(kref_int.ravel(), elems.shape[0]).astype(np.int64)
    return rows, cols, vals


def _finish(name, n_vert, parts, scale, dirichlet):
    rows = np.concatenate([p[0] for p in parts])
    cols = np.concatenate([p[1] for p in parts])
    vals = np.concatenate([p[2] for p in parts])
    touch = dirichlet[rows] | dirichlet[cols]
    vals = np.where(touch, 0, vals)
    # integer sums (exact); explicit zeros stay in the pattern
    A, C, Vint = _coo_to_csr(n_vert, rows, cols, vals.astype(np.int64))
    V = Vint.astype(np.float64) * scale
    return CSRSystem(name, n_vert, A, C, V, dirichlet)


def _p1_element_matrix(coords):
    """Integer (b_i.b_j) matrix of barycentric gradients times (d! * volume),
    for integer vertex coordinates of a simplex of volume 1/d!."""
    coords = np.asarray(coords, dtype=np.float64)
    d = coords.shape[1]
    T = (coords[1:] - coords[0]).T                 # d x d
    Tinv = np.linalg.inv(T)
    G = np.vstack([-Tinv.sum(axis=0), Tinv])       # gradients of barycentrics (rows)
    K = G @ G.T
    Ki = np.rint(K).astype(np.int64)
    assert np.allclose(K, Ki), "reference simplex must have integer gradient products"
    return Ki


def _element_assembled_2d(levels: int) -> CSRSystem:
    """Element-by-element P1 assembly on the refined unit square (slow; the
    structured builder below is checked against it)."""
    m = 2 ** levels
    nx = m + 1
    ix, iy = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    vid = lambda a, b: (b * nx + a)
    v00, v10, v01, v11 = vid(ix, iy), vid(ix + 1, iy), vid(ix, iy + 1), vid(ix + 1, iy + 1)
    # triangles {0,1,3} and {0,2,3} of mesh/examples2D.hpp:205-206
    t1 = np.stack([v00, v10, v11], axis=1)
    t2 = np.stack([v00, v01, v11], axis=1)
    k1 = _p1_element_matrix([[0, 0], [1, 0], [1, 1]])  # units of 1/2 (area 1/2, K = G G^T / 2)
    k2 = _p1_element_matrix([[0, 0], [0, 1], [1, 1]])
    gx, gy = np.meshgrid(np.arange(nx), np.arange(nx), indexing="xy")
    gx, gy = gx.ravel(), gy.ravel()
    dirichlet = (gx == 0) | (gy == 0) | (gx == m) | (gy == m)
    parts = [_assemble(nx * nx, t1, k1, 0.5, dirichlet), _assemble(nx * nx, t2, k2, 0.5, dirichlet)]
    return _finish(f"poisson2d_p1_L{levels}", nx * nx, parts, 0.5, dirichlet)


_KUHN = [(4, 6, 7), (4, 5, 7), (2, 6, 7), (2, 3, 7), (1, 5, 7), (1, 3, 7)]  # mesh/examples3D.hpp:27-32


def _element_assembled_3d(n: int) -> CSRSystem:
    """Element-by-element P1 assembly on n^3 Kuhn cubes (slow; the structured
    builder below is checked against it)."""
    nx = n + 1
    h = 1.0 / n
    ix, iy, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()

    def corner(bits):  # bit 2 = x, bit 1 = y, bit 0 = z (mesh/examples3D.hpp:18-25)
        dx, dy, dz = (bits >> 2) & 1, (bits >> 1) & 1, bits & 1
        return (ix + dx) + nx * ((iy + dy) + nx * (iz + dz))

    parts = []
    dirichlet = _boundary_3d(n)
    for path in _KUHN:
        bits = (0,) + path
        coords = [[(bb >> 2) & 1, (bb >> 1) & 1, bb & 1] for bb in bits]
        kint = _p1_element_matrix(coords)            # K = h * (G G^T) / 6
        elems = np.stack([corner(bb) for bb in bits], axis=1)
        parts.append(_assemble(nx ** 3, elems, kint, h / 6.0, dirichlet))
    return _finish(f"poisson3d_p1_kuhn_{n}", nx ** 3, parts, h / 6.0, dirichlet)


def _boundary_3d(n):
    nx = n + 1
    g = np.arange(nx)
    gz, gy, gx = np.meshgrid(g, g, g, indexing="ij")  # vertex id = gx + nx*(gy + nx*gz)
    return ((gx == 0) | (gy == 0) | (gz == 0) | (gx == n) | (gy == n) | (gz == n)).ravel()


def _structured(name, dims, stencil_units, scale, dirichlet):
    """CSR of a translation-invariant stencil on a vertex grid (x fastest),
    keeping every in-grid neighbour; couplings with a Dirichlet end are 0."""
    dims = tuple(dims)
    nvert = int(np.prod(dims))
    coords = np.stack(np.unravel_index(np.arange(nvert), dims[::-1])[::-1], axis=1)  # (x, y, z)
    strides = np.cumprod((1,) + dims[:-1])
    offs = sorted(stencil_units.items(), key=lambda kv: int(np.dot(kv[0], strides)))
    delta = np.array([o for o, _ in offs])
    units = np.array([u for _, u in offs], dtype=np.int64)
    nb = coords[:, None, :] + delta[None, :, :]
    inside = np.all((nb >= 0) & (nb < np.array(dims)), axis=2)
    cols = (nb * strides).sum(axis=2)
    A = np.zeros(nvert + 1, dtype=np.int64)
    A[1:] = np.cumsum(inside.sum(axis=1))
    C = cols[inside].astype(np.int32)
    rows = np.repeat(np.arange(nvert), inside.sum(axis=1))
    U = np.broadcast_to(units, inside.shape)[inside]
    U = np.where(dirichlet[rows] | dirichlet[C], 0, U)
    return CSRSystem(name, nvert, A.astype(np.int32), C, U.astype(np.float64) * scale, dirichlet)


def _offsets_units(patch: CSRSystem, centre_xyz, dims, scale):
    strides = np.cumprod((1,) + tuple(dims[:-1]))
    centre = int(np.dot(centre_xyz, strides))
    lo, hi = patch.A[centre], patch.A[centre + 1]
    out = {}
    for c, v in zip(patch.C[lo:hi], patch.V[lo:hi]):
        xyz = np.array(np.unravel_index(int(c), tuple(dims)[::-1])[::-1])
        out[tuple(int(t) for t in xyz - np.array(centre_xyz))] = int(round(v / scale))
    return out


def poisson2d_p1(levels: int, numbering: str = "lexicographic") -> CSRSystem:
    """P1 stiffness of UnitSquare2D_simple refined `levels` times: (2^L+1)^2
    vertices.  numbering "reference": the vertices in the order the
    reference's uniform refinement numbers them (reference_permutation)."""
    m = 2 ** levels
    nx = m + 1
    patch = _element_assembled_2d(2)                  # 5x5 vertices, interior stencil at (2,2)
    units = _offsets_units(patch, (2, 2), (5, 5), 0.5)
    gx, gy = np.meshgrid(np.arange(nx), np.arange(nx), indexing="xy")
    dirichlet = ((gx == 0) | (gy == 0) | (gx == m) | (gy == m)).ravel()
    lex = _structured(f"poisson2d_p1_L{levels}", (nx, nx), units, 0.5, dirichlet)
    return _numbered(lex, 2, levels, numbering)


def poisson3d_p1_kuhn(n: int, numbering: str = "lexicographic") -> CSRSystem:
    """P1 stiffness on n^3 cubes of six Kuhn tetrahedra each: (n+1)^3 vertices.
    numbering "reference" (n a power of two): UnitCube3D refined log2(n)
    times, the vertices in the reference's refinement order."""
    nx = n + 1
    patch = _element_assembled_3d(4)                  # 5^3 vertices, interior stencil at (2,2,2)
    units = _offsets_units(patch, (2, 2, 2), (5, 5, 5), (1.0 / 4) / 6.0)
    lex = _structured(f"poisson3d_p1_kuhn_{n}", (nx, nx, nx), units, (1.0 / n) / 6.0,
                      _boundary_3d(n))
    if numbering == "reference" and (n & (n - 1)):
        raise ValueError("the reference's refinement gives 2^L cubes per axis")
    return _numbered(lex, 3, n.bit_length() - 1, numbering)


def _numbered(lex: CSRSystem, dim: int, levels: int, numbering: str) -> CSRSystem:
    if numbering == "lexicographic":
        return lex
    if numbering != "reference":
        raise ValueError(f"unknown numbering {numbering!r}")
    return permute_symmetric(lex, reference_permutation(dim, levels), lex.name + "_refnum")


# ---------------------------------------------------------------------------
# The reference's vertex numbering: uniform refinement of the unit square /
# unit cube as mesh/mesh.simplicial2D.cpp:2205-2750 and
# mesh/mesh.simplicial3D.ur.cxx:134-1005 number it.  A refinement keeps the
# old vertices and appends one vertex per old edge (vertex V+e is the midpoint
# of edge e); the new edges are the two halves of every old edge (e: back
# half, E+e: front half), then three per old face (2E + k*F + f) and, in 3D,
# one per old tetrahedron (2E + 3F + t, the 02-13 diagonal).  Which edge gets
# which index therefore depends on the local order of the edges of every
# face and tetrahedron, which the refinement fixes child by child; the tables
# below restate those assignments.  Level 0 comes from the mesh constructors
# (mesh/mesh.simplicial2D.cpp:56-250, mesh/mesh.simplicial3D.cpp:70-460):
# simplices with sorted vertices, sorted; faces and edges the sorted unique
# lists of their sub-simplices; local edges 01 02 (03) 12 (13 23), local
# faces 012 013 023 123.

def _sorted_unique_rows(a):
    return np.unique(a, axis=0)


def _row_lookup(table, rows):
    """Index of each row of `rows` in the sorted unique `table`."""
    key_t = table @ (1 << (16 * np.arange(table.shape[1])[::-1]))
    key_r = rows @ (1 << (16 * np.arange(rows.shape[1])[::-1]))
    idx = np.searchsorted(key_t, key_r)
    assert np.array_equal(table[idx], rows)
    return idx


def _reference_mesh_2d(levels: int, edges: bool = False):
    """(lattice coordinates in units of 2^-levels, triangles) of
    UnitSquare2D_simple (mesh/examples2D.hpp:194-211) refined `levels` times,
    numbered as the reference numbers them; with edges=True also (edge
    vertex pairs, the three edges of every triangle: 01, 02, 12)."""
    s = 1 << levels
    X = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.int64) * s
    TV = np.sort(np.array([[0, 1, 3], [0, 2, 3]], dtype=np.int64), axis=1)
    TV = _sorted_unique_rows(TV)
    loc = [(0, 1), (0, 2), (1, 2)]
    EV = _sorted_unique_rows(np.concatenate([TV[:, [a, b]] for a, b in loc]))
    TE = np.stack([_row_lookup(EV, TV[:, [a, b]]) for a, b in loc], axis=1)
    for _ in range(levels):
        V, E, T = len(X), len(EV), len(TV)
        X = np.concatenate([X, (X[EV[:, 0]] + X[EV[:, 1]]) // 2])
        mid = V + TE                      # v01, v02, v12 of every triangle
        e_ = np.arange(E)
        t_ = np.arange(T)
        EV = np.concatenate([
            np.stack([EV[:, 0], V + e_], axis=1), np.stack([V + e_, EV[:, 1]], axis=1),
            np.stack([mid[:, 0], mid[:, 1]], axis=1), np.stack([mid[:, 0], mid[:, 2]], axis=1),
            np.stack([mid[:, 1], mid[:, 2]], axis=1)])
        i01_02, i01_12, i02_12 = 2 * E + t_, 2 * E + T + t_, 2 * E + 2 * T + t_
        TE = np.concatenate([
            np.stack([TE[:, 0], TE[:, 1], i01_02], axis=1),
            np.stack([E + TE[:, 0], i01_12, TE[:, 2]], axis=1),
            np.stack([i02_12, E + TE[:, 1], E + TE[:, 2]], axis=1),
            np.stack([i01_02, i01_12, i02_12], axis=1)])
        v00, v11, v22 = TV[:, 0], TV[:, 1], TV[:, 2]
        v01, v02, v12 = mid[:, 0], mid[:, 1], mid[:, 2]
        TV = np.concatenate([np.stack([v00, v01, v02], axis=1), np.stack([v01, v11, v12], axis=1),
                             np.stack([v02, v12, v22], axis=1), np.stack([v01, v02, v12], axis=1)])
    if edges:
        return X, TV, EV, TE
    return X, TV


def _reference_mesh_3d(levels: int):
    """(lattice coordinates in units of 2^-levels, tetrahedra) of UnitCube3D
    (mesh/examples3D.hpp:14-37) refined `levels` times, numbered as the
    reference numbers them."""
    s = 1 << levels
    X = np.array([[(b >> 2) & 1, (b >> 1) & 1, b & 1] for b in range(8)], dtype=np.int64) * s
    TV = np.array([[0, 4, 6, 7], [0, 4, 5, 7], [0, 2, 6, 7], [0, 2, 3, 7], [0, 1, 5, 7],
                   [0, 1, 3, 7]], dtype=np.int64)
    TV = _sorted_unique_rows(np.sort(TV, axis=1))
    floc = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    eloc = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    FV = _sorted_unique_rows(np.concatenate([TV[:, list(f)] for f in floc]))
    EV = _sorted_unique_rows(np.concatenate([TV[:, list(e)] for e in eloc]))
    TE = np.stack([_row_lookup(EV, TV[:, list(e)]) for e in eloc], axis=1)
    TF = np.stack([_row_lookup(FV, TV[:, list(f)]) for f in floc], axis=1)
    FE = np.stack([_row_lookup(EV, FV[:, [a, b]]) for a, b in [(0, 1), (0, 2), (1, 2)]], axis=1)
    st = lambda *cols: np.stack(cols, axis=1)
    for _ in range(levels):
        V, E, F, T = len(X), len(EV), len(FE), len(TV)
        X = np.concatenate([X, (X[EV[:, 0]] + X[EV[:, 1]]) // 2])
        e_, f_, t_ = np.arange(E), np.arange(F), np.arange(T)
        fm = V + FE                       # midpoints of a face's edges 01 02 12
        tm = V + TE                       # midpoints of a tetrahedron's edges 01 02 03 12 13 23
        EV = np.concatenate([
            st(EV[:, 0], V + e_), st(V + e_, EV[:, 1]),
            st(fm[:, 0], fm[:, 1]), st(fm[:, 0], fm[:, 2]), st(fm[:, 1], fm[:, 2]),
            st(tm[:, 1], tm[:, 4])])
        # edges of the new faces and tetrahedra
        fe01_02, fe01_12, fe02_12 = 2 * E + f_, 2 * E + F + f_, 2 * E + 2 * F + f_
        FE_out = np.concatenate([
            st(FE[:, 0], FE[:, 1], fe01_02),
            st(E + FE[:, 0], fe01_12, FE[:, 2]),
            st(fe02_12, E + FE[:, 1], E + FE[:, 2]),
            st(fe01_02, fe01_12, fe02_12)])
        face_e = lambda k, j: 2 * E + k * F + TF[:, j]  # k-th new edge of the tet's old face j
        e01_02, e01_12, e02_12 = face_e(0, 0), face_e(1, 0), face_e(2, 0)
        e01_03, e01_13, e03_13 = face_e(0, 1), face_e(1, 1), face_e(2, 1)
        e02_03, e02_23, e03_23 = face_e(0, 2), face_e(1, 2), face_e(2, 2)
        e12_13, e12_23, e13_23 = face_e(0, 3), face_e(1, 3), face_e(2, 3)
        e02_13 = 2 * E + 3 * F + t_
        FE_in = np.concatenate([
            st(e01_02, e01_03, e02_03), st(e01_12, e01_13, e12_13),
            st(e02_12, e02_23, e12_23), st(e03_13, e03_23, e13_23),
            st(e01_02, e01_13, e02_13), st(e02_03, e02_13, e03_13),
            st(e02_12, e02_13, e12_13), st(e02_13, e02_23, e13_23)])
        e00_01, e01_11 = TE[:, 0], E + TE[:, 0]
        e00_02, e02_22 = TE[:, 1], E + TE[:, 1]
        e00_03, e03_33 = TE[:, 2], E + TE[:, 2]
        e11_12, e12_22 = TE[:, 3], E + TE[:, 3]
        e11_13, e13_33 = TE[:, 4], E + TE[:, 4]
        e22_23, e23_33 = TE[:, 5], E + TE[:, 5]
        TE_new = np.concatenate([
            st(e00_01, e00_02, e00_03, e01_02, e01_03, e02_03),
            st(e01_11, e01_12, e01_13, e11_12, e11_13, e12_13),
            st(e02_12, e02_22, e02_23, e12_22, e12_23, e22_23),
            st(e03_13, e03_23, e03_33, e13_23, e13_33, e23_33),
            st(e01_02, e01_03, e01_13, e02_03, e02_13, e03_13),
            st(e01_02, e01_12, e01_13, e02_12, e02_13, e12_13),
            st(e02_03, e02_13, e02_23, e03_13, e03_23, e13_23),
            st(e02_12, e02_13, e02_23, e12_13, e12_23, e13_23)])
        out_f = lambda k, j: k * F + TF[:, j]           # k-th child of the tet's old face j
        in_f = lambda k: 4 * F + k * T + t_             # k-th interior face of the tet
        TF_new = np.concatenate([
            st(out_f(0, 0), out_f(0, 1), out_f(0, 2), in_f(0)),
            st(out_f(1, 0), out_f(1, 1), in_f(1), out_f(0, 3)),
            st(out_f(2, 0), in_f(2), out_f(1, 2), out_f(1, 3)),
            st(in_f(3), out_f(2, 1), out_f(2, 2), out_f(2, 3)),
            st(in_f(0), in_f(4), out_f(3, 1), in_f(5)),
            st(out_f(3, 0), in_f(4), in_f(1), in_f(6)),
            st(in_f(5), out_f(3, 2), in_f(7), in_f(3)),
            st(in_f(6), in_f(2), in_f(7), out_f(3, 3))])
        v00, v11, v22, v33 = TV[:, 0], TV[:, 1], TV[:, 2], TV[:, 3]
        v01, v02, v03, v12, v13, v23 = (tm[:, k] for k in range(6))
        TV = np.concatenate([
            st(v00, v01, v02, v03), st(v01, v11, v12, v13), st(v02, v12, v22, v23),
            st(v03, v13, v23, v33), st(v01, v02, v03, v13), st(v01, v02, v12, v13),
            st(v02, v03, v13, v23), st(v02, v12, v13, v23)])
        FE = np.concatenate([FE_out, FE_in])
        TE, TF = TE_new, TF_new
    return X, TV


def reference_permutation(dim: int, levels: int) -> np.ndarray:
    """perm[i] = lexicographic (x fastest) index of the reference's vertex i
    on the (2^levels+1)^dim grid."""
    X, _ = (_reference_mesh_2d if dim == 2 else _reference_mesh_3d)(levels)
    nx = (1 << levels) + 1
    strides = nx ** np.arange(dim)
    return (X * strides).sum(axis=1)


def permute_symmetric(system: CSRSystem, perm: np.ndarray, name: str) -> CSRSystem:
    """P^T A P with row i of the result = row perm[i] of A (columns renumbered,
    each row sorted by column, as MatrixCSR stores them)."""
    n = system.n
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(system.A))
    new_rows = inv[rows]
    new_cols = inv[system.C]
    order = np.lexsort((new_cols, new_rows))
    A = np.zeros(n + 1, dtype=np.int64)
    np.add.at(A, new_rows + 1, 1)
    A = np.cumsum(A)
    return CSRSystem(name, n, A.astype(np.int32), new_cols[order].astype(np.int32),
                     system.V[order].copy(), system.dirichlet[perm].copy())


def splitmix64(seed: int, count: int) -> np.ndarray:
    """SplitMix64 outputs for counters seed+1 .. seed+count (uint64)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.arange(1, count + 1, dtype=np.uint64)
             * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed: int, count: int) -> np.ndarray:
    """U[-1,1) doubles: 53 high bits of SplitMix64."""
    u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return 2.0 * u - 1.0


def rhs_uniform(system: CSRSystem, seed: int) -> np.ndarray:
    b = uniform_pm1(seed, system.n)
    b[system.dirichlet] = 0.0
    return b


def random_spd(n: int, density: float, seed: int) -> CSRSystem:
    """Random sparse symmetric diagonally dominant matrix (positive definite)."""
    rng = np.random.default_rng(seed)
    m = max(1, int(density * n * n / 2))
    i = rng.integers(0, n, m)
    j = rng.integers(0, n, m)
    keep = i != j
    i, j = i[keep], j[keep]
    v = rng.uniform(-1, 1, i.shape[0])
    rows = np.concatenate([i, j])
    cols = np.concatenate([j, i])
    vals = np.concatenate([v, v])
    A0, C0, V0 = _coo_to_csr(n, rows, cols, vals)
    rowsum = np.zeros(n)
    np.add.at(rowsum, np.repeat(np.arange(n), np.diff(A0)), np.abs(V0))
    rows = np.concatenate([np.repeat(np.arange(n), np.diff(A0)), np.arange(n)])
    cols = np.concatenate([C0, np.arange(n)])
    vals = np.concatenate([V0, rowsum + rng.uniform(0.5, 1.5, n)])
    A, C, V = _coo_to_csr(n, rows, cols, vals)
    return CSRSystem(f"random_spd_{n}_{seed}", n, A, C, V, np.zeros(n, dtype=bool))


def _csr_from_int_coo(n, rows, cols, vals_list):
    """COO with several integer value arrays -> CSR (sorted columns), exact
    integer duplicate sums; returns A, C and one summed array per input."""
    key = rows.astype(np.int64) * n + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key = key[order]
    start = np.concatenate([[True], key[1:] != key[:-1]])
    idx = np.nonzero(start)[0]
    uniq = key[idx]
    sums = [np.add.reduceat(v[order].astype(np.int64), idx) for v in vals_list]
    r = uniq // n
    A = np.zeros(n + 1, dtype=np.int64)
    np.add.at(A, r + 1, 1)
    return np.cumsum(A).astype(np.int32), (uniq % n).astype(np.int32), sums


def _whitney1_element(coords):
    """Integer curl-curl and mass matrices of the six Whitney 1-forms of a
    unit Kuhn tetrahedron (volume 1/6), edge (i, j), i < j local:
      curl-curl: K_ef = vol*4*(g_i x g_j).(g_k x g_l)  = (2/3) * Kint
      mass:      M_ef = vol/20 * sum (1+d_ac) g_b.g_d ... = (1/120) * Mint
    with g = gradients of the barycentric coordinates."""
    coords = np.asarray(coords, dtype=np.float64)
    T = (coords[1:] - coords[0]).T
    Tinv = np.linalg.inv(T)
    g = np.vstack([-Tinv.sum(axis=0), Tinv])
    edges = [(i, j) for i in range(4) for j in range(i + 1, 4)]
    K = np.zeros((6, 6))
    M = np.zeros((6, 6))
    dl = lambda a, b: 1.0 if a == b else 0.0
    for e, (i, j) in enumerate(edges):
        for f, (k, l) in enumerate(edges):
            K[e, f] = 4.0 * np.dot(np.cross(g[i], g[j]), np.cross(g[k], g[l])) / 6.0
            M[e, f] = ((1 + dl(i, k)) * g[j] @ g[l] - (1 + dl(i, l)) * g[j] @ g[k]
                       - (1 + dl(j, k)) * g[i] @ g[l] + (1 + dl(j, l)) * g[i] @ g[k]) / 120.0
    Kint = np.rint(K * 1.5).astype(np.int64)
    Mint = np.rint(M * 120.0).astype(np.int64)
    assert np.allclose(K * 1.5, Kint) and np.allclose(M * 120.0, Mint)
    return edges, Kint, Mint


def hcurl3d_whitney_kuhn(n: int, curl_weight: float = 1.0, mass_weight: float = 1.0,
                         boundary: bool = True) -> CSRSystem:
    """curl-curl + mass of lowest-order Whitney 1-forms (edge elements) on n^3
    Kuhn cubes, edges oriented from the lower to the higher vertex number;
    edges lying in the boundary carry the tangential Dirichlet condition
    (their rows and columns are 0, structure kept).  For n = 48: 795,024 edges."""
    nx = n + 1
    h = 1.0 / n
    ix, iy, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()

    def corner(bits):
        dx, dy, dz = (bits >> 2) & 1, (bits >> 1) & 1, bits & 1
        return (ix + dx) + nx * ((iy + dy) + nx * (iz + dz))

    tet_edges, tet_K, tet_M = [], [], []
    for path in _KUHN:
        bits = (0,) + path
        coords = [[(bb >> 2) & 1, (bb >> 1) & 1, bb & 1] for bb in bits]
        edges, Kint, Mint = _whitney1_element(coords)
        verts = np.stack([corner(bb) for bb in bits], axis=1)  # (cubes, 4), increasing ids
        ga = np.stack([verts[:, i] for i, j in edges], axis=1)
        gb = np.stack([verts[:, j] for i, j in edges], axis=1)
        assert (ga < gb).all()  # a Kuhn path visits vertices in increasing id order
        tet_edges.append(ga.astype(np.int64) * nx ** 3 + gb)
        tet_K.append(Kint)
        tet_M.append(Mint)
    allkeys = np.concatenate([t.ravel() for t in tet_edges])
    ekeys, inverse = np.unique(allkeys, return_inverse=True)
    ne = ekeys.shape[0]
    rows, cols, kv, mv = [], [], [], []
    off = 0
    for t in range(len(_KUHN)):
        cnt = tet_edges[t].size
        eid = inverse[off:off + cnt].reshape(-1, 6)
        off += cnt
        rows.append(np.repeat(eid, 6, axis=1).ravel())
        cols.append(np.tile(eid, (1, 6)).ravel())
        kv.append(np.tile(tet_K[t].ravel(), eid.shape[0]))
        mv.append(np.tile(tet_M[t].ravel(), eid.shape[0]))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    kv, mv = np.concatenate(kv), np.concatenate(mv)
    # boundary edges: both ends on one boundary plane
    a, b = ekeys // nx ** 3, ekeys % nx ** 3
    ca = np.stack([a % nx, (a // nx) % nx, a // nx ** 2], axis=1)
    cb = np.stack([b % nx, (b // nx) % nx, b // nx ** 2], axis=1)
    dirichlet = (((ca == 0) & (cb == 0)) | ((ca == n) & (cb == n))).any(axis=1)
    if not boundary:
        dirichlet = np.zeros(ne, dtype=bool)
    touch = dirichlet[rows] | dirichlet[cols]
    kv = np.where(touch, 0, kv)
    mv = np.where(touch, 0, mv)
    A, C, (Ks, Ms) = _csr_from_int_coo(ne, rows, cols, [kv, mv])
    V = (curl_weight * (2.0 / (3.0 * h))) * Ks.astype(np.float64) \
        + (mass_weight * (h / 120.0)) * Ms.astype(np.float64)
    system = CSRSystem(f"hcurl3d_whitney_kuhn_{n}", ne, A, C, V, dirichlet)
    system.edges = np.stack([a, b], axis=1)  # vertex pair of each edge
    return system


# --------------------------------------------------------------------------- P_r Lagrange

def _poly_mul(p, q):
    out = {}
    for ea, ca in p.items():
        for eb, cb in q.items():
            e = tuple(x + y for x, y in zip(ea, eb))
            out[e] = out.get(e, 0.0) + ca * cb
    return out


def _lagrange_barycentric(d, r):
    """Principal-lattice Lagrange basis of degree r on a d-simplex as
    polynomials in the d+1 barycentric coordinates:
    phi_a = prod_i prod_{k<a_i} (r*lam_i - k)/(k+1),  |a| = r."""
    import itertools

    alphas = [a for a in itertools.product(range(r + 1), repeat=d + 1) if sum(a) == r]
    basis = []
    for a in alphas:
        p = {tuple([0] * (d + 1)): 1.0}
        for i in range(d + 1):
            for k in range(a[i]):
                e1 = tuple(1 if j == i else 0 for j in range(d + 1))
                p = _poly_mul(p, {e1: r / (k + 1.0), tuple([0] * (d + 1)): -k / (k + 1.0)})
        basis.append({e: c for e, c in p.items() if c != 0.0})
    return alphas, basis


def _pr_element_matrix(coords, r):
    """Stiffness of the degree-r Lagrange basis on one simplex with integer
    vertex coordinates: K_ij = int grad phi_i . grad phi_j, exact monomial
    integration int lam^b = d! vol b! / (|b| + d)!."""
    from math import factorial

    coords = np.asarray(coords, dtype=np.float64)
    d = coords.shape[1]
    T = (coords[1:] - coords[0]).T
    vol = abs(np.linalg.det(T)) / factorial(d)
    Tinv = np.linalg.inv(T)
    G = np.vstack([-Tinv.sum(axis=0), Tinv])  # grad lam_i
    alphas, basis = _lagrange_barycentric(d, r)

    def deriv(p, i):
        out = {}
        for e, c in p.items():
            if e[i] > 0:
                e2 = tuple(x - (1 if j == i else 0) for j, x in enumerate(e))
                out[e2] = out.get(e2, 0.0) + c * e[i]
        return out

    def integral(p):
        tot = 0.0
        for e, c in p.items():
            num = np.prod([factorial(x) for x in e])
            tot += c * factorial(d) * vol * num / factorial(sum(e) + d)
        return tot

    dphi = [[deriv(p, i) for i in range(d + 1)] for p in basis]
    nb = len(basis)
    K = np.zeros((nb, nb))
    GG = G @ G.T
    for a in range(nb):
        for b in range(a, nb):
            s = 0.0
            for i in range(d + 1):
                for j in range(d + 1):
                    if GG[i, j] != 0.0 and dphi[a][i] and dphi[b][j]:
                        s += GG[i, j] * integral(_poly_mul(dphi[a][i], dphi[b][j]))
            K[a, b] = K[b, a] = s
    return alphas, K


_PR_CACHE = {}


def lagrange_pr(dim: int, n: int, r: int) -> CSRSystem:
    """Degree-r Lagrange stiffness on the unit square (two triangles per cell,
    mesh/examples2D.hpp:194-211) or unit cube (six Kuhn tetrahedra per cell,
    mesh/examples3D.hpp:14-37) with n cells per axis: (r*n+1)^dim nodes on the
    principal lattice, numbered lexicographically; boundary nodes Dirichlet.
    For r = 1 this is poisson2d_p1 / poisson3d_p1_kuhn up to value rounding."""
    cache_key = (dim, n, r)
    if cache_key in _PR_CACHE:
        return _PR_CACHE[cache_key]
    m = r * n + 1
    if dim == 2:
        simplices = [[(0, 0), (1, 0), (1, 1)], [(0, 0), (0, 1), (1, 1)]]
    else:
        simplices = [[((bb >> 2) & 1, (bb >> 1) & 1, bb & 1) for bb in (0,) + path] for path in _KUHN]
    strides = np.array([m ** k for k in range(dim)])
    N = m ** dim
    grid = np.stack(np.unravel_index(np.arange(N), (m,) * dim)[::-1], axis=1)
    dirichlet = ((grid == 0) | (grid == m - 1)).any(axis=1)
    elems = []
    for simp in simplices:
        alphas, K = _pr_element_matrix(simp, r)
        V0 = np.array(simp)
        local = np.array([r * V0[0] + sum(a[i] * (V0[i] - V0[0]) for i in range(1, dim + 1))
                          for a in alphas])  # lattice coordinates of the nodes of one cell
        elems.append((local, (local * strides).sum(axis=1), K))
    # Every entry couples two nodes of one simplex, so col - row is one of a
    # few lattice offsets: accumulate into an (offset, node) array.  For one
    # simplex type and one local pair (a, b), node a of every cell is a
    # strided sub-lattice and the offset is fixed, so each update is a
    # strided slice; contributions to an entry add in (simplex type, a) order.
    offsets = np.unique(np.concatenate([(ll[None, :] - ll[:, None]).ravel() for _, ll, _ in elems]))
    u = len(offsets)
    W = np.zeros((u,) + (m,) * dim)      # axes: offset, coordinate dim-1 .. 0
    P = np.zeros((u,) + (m,) * dim, dtype=bool)
    D = dirichlet.reshape((m,) * dim)
    scale = (1.0 / n) ** (dim - 2)

    def sub(c):  # node with cell-local lattice coordinates c, in every cell
        return tuple(slice(c[i], c[i] + r * n, r) for i in reversed(range(dim)))

    for local, ll, K in elems:
        for a in range(len(ll)):
            sa = sub(local[a])
            da = D[sa]
            ks = np.searchsorted(offsets, ll - ll[a])
            for b in range(len(ll)):
                W[(ks[b],) + sa] += np.where(da | D[sub(local[b])], 0.0, K[a, b]) * scale
                P[(ks[b],) + sa] = True
    P = np.ascontiguousarray(P.reshape(u, N).T)                       # (node, offset)
    row, k = np.nonzero(P)                                            # row-major, offsets ascending
    V = W.reshape(-1)[k * N + row]
    del W
    A = np.zeros(N + 1, dtype=np.int64)
    A[1:] = P.sum(axis=1)
    del P
    out = CSRSystem(f"lagrange{dim}d_n{n}_r{r}", N, np.cumsum(A).astype(np.int32),
                    (row + offsets[k]).astype(np.int32), V, dirichlet)
    _PR_CACHE[cache_key] = out
    return out


def _multi_indices(m: int, degree: int):
    """generateMultiIndices over m entries of total `degree`, in the
    reference's order (combinatorics/generatemultiindices.cpp:17-72): the
    candidates 0, 1, 2, ... written in base degree+1 with entry 0 as the least
    significant digit, those whose digits sum to `degree`."""
    base = degree + 1
    out = []
    for cand in range(degree, base ** m):
        digits, c = [], cand
        for _ in range(m):
            digits.append(c % base)
            c //= base
        if sum(digits) == degree:
            out.append(tuple(digits))
    return out


def _seq_segment_sum(keys, vals):
    """Unique keys (ascending) and, per key, the sum of its values added one
    after another in the order given (a sequential left-to-right sum)."""
    order = np.argsort(keys, kind="stable")
    k, v = keys[order], vals[order]
    uk, start, count = np.unique(k, return_index=True, return_counts=True)
    acc = v[start].copy()
    for t in range(1, int(count.max()) if len(count) else 0):
        has = count > t
        acc[has] += v[start[has] + t]
    return uk, acc, count


_SULLIVAN_CACHE = {}


def sullivan_pr_2d(levels: int, r: int) -> CSRSystem:
    """The reference's degree-r Lagrange stiffness in FEEC's Sullivan basis on
    UnitSquare2D_simple refined `levels` times with automatic Dirichlet flags,
    as tests/sullivan2D/poissondirichlet2D.cpp:120-151 composes it:
    K = (D I)^T M_1 (D I) with the inclusion I = FEECSullivanInclusionMatrix
    (fem/global.inclsullivan.cpp:19-196), the broken differential
    D = FEECBrokenDiffMatrix (fem/global.diffmatrix.cpp:15-115) and the broken
    mass matrix of degree r-1 one-forms M_1 = FEECBrokenMassMatrix
    (fem/global.massmatrix.cpp:20-289), each product a SparseMatrixMultiplication
    (sparse/sparsematrix.cpp:609-722: zeros dropped from both factors, products
    of the shared index summed, zeros dropped again from the factors of the
    next product), then MatrixCSR.

    Unknowns: the vertices, then (r-1) per edge, then (r-1)(r-2)/2 per
    triangle, each in the mesh's own numbering; a basis function is a
    barycentric monomial lambda^alpha supported on the face alpha spans.  Rows
    and columns of Dirichlet unknowns have no entries (the inclusion's entries
    for Dirichlet faces are zero), so their diagonal is 0.  Sums of several
    products run in the order of the shared index (the reference's merge
    order; its final std::sort may reorder equal keys, which moves a value by
    a few units in the last place at most)."""
    key = (levels, r)
    if key in _SULLIVAN_CACHE:
        return _SULLIVAN_CACHE[key]
    if not 1 <= r <= 6:
        raise ValueError("degree r must be 1..6")
    X, TV, EV, TE = _reference_mesh_2d(levels, edges=True)
    nv, ne, nt = len(X), len(EV), len(TV)
    # automatic_dirichlet_flags (mesh/mesh.cpp:340-363): edges of one triangle,
    # and their vertices
    edge_dir = np.bincount(TE.ravel(), minlength=ne) == 1
    vert_dir = np.zeros(nv, dtype=bool)
    vert_dir[EV[edge_dir].ravel()] = True

    # continuous unknowns: ListOfSullivanIndices(d, 0, r) (fem/indexfunctions.cpp:15-84)
    lists = [[a for a in _multi_indices(d + 1, r) if min(a) >= 1] for d in range(3)]
    offset = [0, nv * len(lists[0]), nv * len(lists[0]) + ne * len(lists[1])]
    n = offset[2] + nt * len(lists[2])

    # the local broken space of degree r and the one-forms of degree r-1
    broken = _multi_indices(3, r)
    bidx = {a: i for i, a in enumerate(broken)}
    low = _multi_indices(3, r - 1)
    lidx = {a: i for i, a in enumerate(low)}
    nb, nf = len(broken), 3 * len(low)

    # inclusion: local broken function -> continuous unknown (-1: Dirichlet)
    glob = np.full((nt, nb), -1, dtype=np.int64)
    s_idx = np.arange(nt)
    for fi in range(3):  # vertices
        vol = [0, 0, 0]
        vol[fi] = r
        glob[:, bidx[tuple(vol)]] = np.where(vert_dir[TV[:, fi]], -1, offset[0] + TV[:, fi])
    for fi, (a, b) in enumerate(((0, 1), (0, 2), (1, 2))):  # edges: TE[:, fi] joins a, b
        e = TE[:, fi]
        for ai, alpha in enumerate(lists[1]):
            vol = [0, 0, 0]
            vol[a], vol[b] = alpha
            glob[:, bidx[tuple(vol)]] = np.where(edge_dir[e], -1,
                                                 offset[1] + e * len(lists[1]) + ai)
    for ai, alpha in enumerate(lists[2]):  # the triangle
        glob[:, bidx[alpha]] = offset[2] + s_idx * len(lists[2]) + ai

    # broken differential: d lambda^alpha = sum_p alpha_p lambda^(alpha - e_p) d lambda_p
    Dloc = np.zeros((nf, nb))
    for j, a in enumerate(broken):
        for p in range(3):
            if a[p] > 0:
                new = list(a)
                new[p] -= 1
                Dloc[lidx[tuple(new)] * 3 + p, j] = a[p]

    # per triangle: gradient products of the barycentric coordinates and area
    # (mesh/mesh.cpp:525-542,971-1001), exact on this dyadic mesh
    h = 1.0 / (1 << levels)
    P = X[TV].astype(np.float64) * h                       # (nt, 3 vertices, 2)
    J = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]], axis=2)  # columns v1-v0, v2-v0
    JtJ = np.einsum("tki,tkj->tij", J, J)
    det = JtJ[:, 0, 0] * JtJ[:, 1, 1] - JtJ[:, 0, 1] * JtJ[:, 1, 0]
    inv = np.stack([np.stack([JtJ[:, 1, 1], -JtJ[:, 0, 1]], 1),
                    np.stack([-JtJ[:, 1, 0], JtJ[:, 0, 0]], 1)], 1) / det[:, None, None]
    Mul = np.array([[-1.0, 1.0, 0.0], [-1.0, 0.0, 1.0]])
    G = np.einsum("ia,tij,jb->tab", Mul, inv, Mul)         # (nt, 3, 3)
    area = np.abs(J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]) / 2.0

    # polynomialmassmatrix(2, r-1) (fem/polynomialmassmatrix.cpp:15-60)
    def fact(k):
        out = 1.0
        for i in range(2, k + 1):
            out *= i
        return out

    polyMM = np.array([[(2.0 * np.prod([fact(x + y) for x, y in zip(a, b)])) /
                        fact(2 + 2 * (r - 1)) for b in low] for a in low])
    # M_1 = (polyMM (x) G) * area, entry (ip*3+if, jp*3+jf)
    M1 = (polyMM[None, :, None, :, None] * G[:, None, :, None, :]).reshape(nt, nf, nf)
    M1 = M1 * area[:, None, None]

    # opr = D I: broken one-form row i, continuous column glob[j], value D[i, j]
    # (one inclusion entry per broken function, so nothing is summed)
    valid = glob >= 0                                      # (nt, nb)
    # X = M_1 & opr: X[i, col(j)] = sum_k M_1[i, k] * D[k, j] over k ascending,
    # nonzero factors only; `has` marks entries with at least one product
    Xv = np.zeros((nt, nf, nb))
    has = np.zeros((nt, nf, nb), dtype=bool)
    for k in range(nf):
        dk = Dloc[k]                                       # (nb,)
        mk = M1[:, :, k]                                   # (nt, nf)
        use = (mk != 0.0)[:, :, None] & (dk != 0.0)[None, None, :] & valid[:, None, :]
        Xv = np.where(use, Xv + mk[:, :, None] * dk[None, None, :], Xv)
        has |= use
    keep = has & (Xv != 0.0)                               # dropped again as a factor

    # K = opr^T & X: K[col(j), col(l)] = sum over one-form rows i (triangle,
    # then local row ascending) of D[i, j] * X[i, l]
    rows_l, cols_l, vals_l, ord_l = [], [], [], []
    for i in range(nf):
        for j in range(nb):
            if Dloc[i, j] == 0.0:
                continue
            sel = valid[:, j][:, None] & keep[:, i, :]     # (nt, nb)
            t, l = np.nonzero(sel)
            rows_l.append(glob[t, j])
            cols_l.append(glob[t, l])
            vals_l.append(Dloc[i, j] * Xv[t, i, l])
            ord_l.append(t * nf + i)                       # the shared index
    rows = np.concatenate(rows_l)
    cols = np.concatenate(cols_l)
    vals = np.concatenate(vals_l)
    order = np.argsort(np.concatenate(ord_l), kind="stable")
    rows, cols, vals = rows[order], cols[order], vals[order]
    keys, sums, _ = _seq_segment_sum(rows * n + cols, vals)
    Rk, Ck = keys // n, keys % n
    A = np.zeros(n + 1, dtype=np.int64)
    np.add.at(A, Rk + 1, 1)
    dirichlet = np.zeros(n, dtype=bool)
    dirichlet[offset[0]:offset[1]] = vert_dir
    dirichlet[offset[1]:offset[2]] = np.repeat(edge_dir, len(lists[1]))
    out = CSRSystem(f"sullivan2d_L{levels}_r{r}", n, np.cumsum(A).astype(np.int32),
                    Ck.astype(np.int32), sums, dirichlet)
    _SULLIVAN_CACHE[key] = out
    return out


@dataclass
class BlockSystem:
    """Blocks of a saddle-point system [A Bt; B -C] for the Hodge solver
    (solver/systemsparsesolver.hpp:58-72): A is L x L (SPD), B is N x L, Bt
    its transpose, C is N x N; each a (rows, cols, vals) CSR triple."""
    name: str
    N: int
    L: int
    A: tuple
    B: tuple
    Bt: tuple
    C: tuple

    def dense(self):
        """(A, B, Bt, C) as dense arrays."""
        def d(m, r, c):
            M = np.zeros((r, c))
            rows = np.repeat(np.arange(r), np.diff(m[0]))
            np.add.at(M, (rows, m[1]), m[2])
            return M
        return (d(self.A, self.L, self.L), d(self.B, self.N, self.L),
                d(self.Bt, self.L, self.N), d(self.C, self.N, self.N))


def _rect_csr(nrows, ncols, rows, cols, vals):
    import scipy.sparse as sp

    m = sp.coo_matrix((vals, (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64))),
                      shape=(nrows, ncols)).tocsr()
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float64)


def schur_blocks(m: int, coarse: int = 2, c_weight: float = 0.0, seed: int = 0) -> BlockSystem:
    """A model block system on an m x m grid: A = the 5-point stencil
    (4, -1, -1, -1, -1) plus the identity, B = averages over coarse x coarse
    cells of grid points with seeded weights in [1, 2) (disjoint supports,
    full row rank), Bt = B^T, C = c_weight times the 5-point stencil on the
    coarse grid (no entries at all when c_weight is 0, as the reference's
    Darcy tests pass a zero matrix)."""
    assert m % coarse == 0
    lap = fd_laplacian_2d(m)
    L = m * m
    rows = np.repeat(np.arange(L), np.diff(lap.A))
    h2 = 1.0 / (m + 1) ** 2
    A = _rect_csr(L, L, np.concatenate([rows, np.arange(L)]), np.concatenate([lap.C, np.arange(L)]),
                  np.concatenate([lap.V * h2, np.ones(L)]))
    mc = m // coarse
    N = mc * mc
    w = 1.0 + 0.5 * (uniform_pm1(seed, L) + 1.0)
    fine = np.arange(L)
    fx, fy = divmod(fine, m)
    cell = (fx // coarse) * mc + (fy // coarse)
    B = _rect_csr(N, L, cell, fine, w / (coarse * coarse))
    Bt = _rect_csr(L, N, fine, cell, w / (coarse * coarse))
    if c_weight > 0:
        cl = fd_laplacian_2d(mc)
        crows = np.repeat(np.arange(N), np.diff(cl.A))
        C = _rect_csr(N, N, crows, cl.C, c_weight * cl.V / (mc + 1) ** 2)
    else:
        C = (np.zeros(N + 1, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0))
    return BlockSystem(f"schur_m{m}_c{coarse}_w{c_weight:g}", N, L, A, B, Bt, C)
