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

Vertices are numbered lexicographically (x fastest).  The reference numbers
the vertices of a refined mesh in refinement order instead; the matrices are
the same up to that symmetric permutation.

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


def poisson2d_p1(levels: int) -> CSRSystem:
    """P1 stiffness of UnitSquare2D_simple refined `levels` times: (2^L+1)^2 vertices."""
    m = 2 ** levels
    nx = m + 1
    patch = _element_assembled_2d(2)                  # 5x5 vertices, interior stencil at (2,2)
    units = _offsets_units(patch, (2, 2), (5, 5), 0.5)
    gx, gy = np.meshgrid(np.arange(nx), np.arange(nx), indexing="xy")
    dirichlet = ((gx == 0) | (gy == 0) | (gx == m) | (gy == m)).ravel()
    return _structured(f"poisson2d_p1_L{levels}", (nx, nx), units, 0.5, dirichlet)


def poisson3d_p1_kuhn(n: int) -> CSRSystem:
    """P1 stiffness on n^3 cubes of six Kuhn tetrahedra each: (n+1)^3 vertices."""
    nx = n + 1
    patch = _element_assembled_3d(4)                  # 5^3 vertices, interior stencil at (2,2,2)
    units = _offsets_units(patch, (2, 2, 2), (5, 5, 5), (1.0 / 4) / 6.0)
    return _structured(f"poisson3d_p1_kuhn_{n}", (nx, nx, nx), units, (1.0 / n) / 6.0,
                       _boundary_3d(n))


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
