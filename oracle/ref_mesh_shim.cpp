// ref_mesh_shim.cpp -- extern "C" entry points into the reference's own mesh
// refinement and P1 assembly, compiled from the reference sources by
// oracle/Makefile into oracle/_ref/libfeecref.so.
//
// TEST INFRASTRUCTURE ONLY: tests/golden/make_golden.py and tests/ use it to
// pin the build's reference-numbered mesh generator (vertex numbering,
// simplices, CSR structure and values).  The shim adds nothing to the
// algorithms: it builds the reference's unit square / unit cube, refines it,
// sets the Dirichlet flags and assembles the P1 stiffness matrices the
// reference's tests use, and the lowest-order Whitney 1-form curl-curl and
// mass matrices as its 3D curl-curl tests form them
// (tests/sullivan3D/curlcurl3D.cpp:241-312, with the Whitney inclusion of
// tests/fem/condition3Ddense.cpp:220-223 in place of the Sullivan one).
#include <cstring>
#include <vector>

#include "base/include.hpp"
#include "dense/densematrix.hpp"
#include "fem/global.diffmatrix.hpp"
#include "fem/global.inclsullivan.hpp"
#include "fem/global.inclwhitney.hpp"
#include "fem/global.massmatrix.hpp"
#include "fem/lagrangematrices.hpp"
#include "mesh/examples2D.hpp"
#include "mesh/examples3D.hpp"
#include "mesh/mesh.simplicial2D.hpp"
#include "mesh/mesh.simplicial3D.hpp"
#include "operators/floatvector.hpp"
#include "sparse/matcsr.hpp"
#include "sparse/sparsematrix.hpp"

namespace {

struct MeshResult {
    int dim = 0, nv = 0, ns = 0;
    std::vector<double> coords;  // nv x dim
    std::vector<int> simplices;  // ns x (dim+1), the mesh's own vertex order
    std::vector<int> dirichlet;  // per vertex
    // structural P1 stiffness (every vertex pair that shares an edge, explicit
    // zeros kept): LagrangeCoefficientStiffnessMatrix with the identity weight
    std::vector<int> bA, bC;
    std::vector<double> bV;
    // P1 stiffness with exact zeros dropped: MatrixCSR(LagrangeStiffnessMatrix)
    // followed by compressentries, as tests/sullivan2D/poissondirichletlshaped.cpp:128-131
    std::vector<int> aA, aC;
    std::vector<double> aV;
} g_mesh;

void keep_csr(const MatrixCSR& m, std::vector<int>& A, std::vector<int>& C, std::vector<double>& V)
{
    const int n = m.getdimout();
    const int nnz = m.getA()[n];
    A.assign(m.getA(), m.getA() + n + 1);
    C.assign(m.getC(), m.getC() + nnz);
    V.assign(m.getV(), m.getV() + nnz);
}

template <class M>
void collect(M& mesh, int dim, int with_matrices)
{
    MeshResult& r = g_mesh;
    r = MeshResult();
    r.dim = dim;
    r.nv = mesh.count_simplices(0);
    r.ns = mesh.count_simplices(dim);
    r.coords.resize((size_t)r.nv * dim);
    for (int v = 0; v < r.nv; v++)
        for (int d = 0; d < dim; d++) r.coords[(size_t)v * dim + d] = mesh.getCoordinates().getdata(v, d);
    r.simplices.resize((size_t)r.ns * (dim + 1));
    for (int s = 0; s < r.ns; s++)
        for (int k = 0; k <= dim; k++)
            r.simplices[(size_t)s * (dim + 1) + k] = mesh.get_subsimplex(dim, 0, s, k);
    mesh.automatic_dirichlet_flags();
    r.dirichlet.resize(r.nv);
    for (int v = 0; v < r.nv; v++)
        r.dirichlet[v] = mesh.get_flag(0, v) == SimplexFlag::SimplexFlagDirichlet ? 1 : 0;
    if (!with_matrices) return;
    const MatrixCSR b = LagrangeCoefficientStiffnessMatrix(
        mesh, 1, 1, [dim](const FloatVector&) -> DenseMatrix { return IdentityMatrix(dim); });
    keep_csr(b, r.bA, r.bC, r.bV);
    MatrixCSR a(LagrangeStiffnessMatrix(mesh, 1));
    a.compressentries();
    keep_csr(a, r.aA, r.aC, r.aV);
}

struct WhitneyResult {
    int ne = 0;
    std::vector<int> edges;      // ne x 2, the mesh's own vertex order
    std::vector<int> dirichlet;  // per edge
    std::vector<int> kA, kC, mA, mC;
    std::vector<double> kV, mV;
} g_whitney;

}  // namespace

extern "C" {

// Builds UnitSquare2D_simple (dim 2) or UnitCube3D (dim 3), refines it
// `levels` times and keeps its data; with_matrices also assembles the two
// stiffness matrices.  Returns the vertex count.
int feecref_mesh(int dim, int levels, int with_matrices)
{
    if (dim == 2) {
        MeshSimplicial2D M = UnitSquare2D_simple();
        for (int l = 0; l < levels; l++) M.uniformrefinement();
        collect(M, 2, with_matrices);
    } else {
        MeshSimplicial3D M = UnitCube3D();
        for (int l = 0; l < levels; l++) M.uniformrefinement();
        collect(M, 3, with_matrices);
    }
    return g_mesh.nv;
}

// sizes: nv, ns, structural nnz, compressed nnz
void feecref_mesh_sizes(int* out)
{
    out[0] = g_mesh.nv;
    out[1] = g_mesh.ns;
    out[2] = g_mesh.bA.empty() ? 0 : g_mesh.bA.back();
    out[3] = g_mesh.aA.empty() ? 0 : g_mesh.aA.back();
}

void feecref_mesh_copy(double* coords, int* simplices, int* dirichlet, int* bA, int* bC, double* bV,
                       int* aA, int* aC, double* aV)
{
    const MeshResult& r = g_mesh;
    std::memcpy(coords, r.coords.data(), r.coords.size() * sizeof(double));
    std::memcpy(simplices, r.simplices.data(), r.simplices.size() * sizeof(int));
    std::memcpy(dirichlet, r.dirichlet.data(), r.dirichlet.size() * sizeof(int));
    if (!r.bA.empty()) {
        std::memcpy(bA, r.bA.data(), r.bA.size() * sizeof(int));
        std::memcpy(bC, r.bC.data(), r.bC.size() * sizeof(int));
        std::memcpy(bV, r.bV.data(), r.bV.size() * sizeof(double));
        std::memcpy(aA, r.aA.data(), r.aA.size() * sizeof(int));
        std::memcpy(aC, r.aC.data(), r.aC.size() * sizeof(int));
        std::memcpy(aV, r.aV.data(), r.aV.size() * sizeof(double));
    }
}

// UnitCube3D refined `levels` times: curl-curl K = Conjugation(M_2, D & I)
// and mass M = Conjugation(M_1, I) of the Whitney 1-forms, where I is
// FEECWhitneyInclusionMatrix(., 3, 1, 1) (Whitney forms into broken
// P_1 Lambda^1), D = FEECBrokenDiffMatrix(., 3, 1, 1) and M_k the broken mass
// matrices of P_1 Lambda^1 and P_0 Lambda^2.  Keeps the edges' vertices and
// Dirichlet flags (automatic_dirichlet_flags); returns the edge count.
int feecref_whitney1(int levels)
{
    MeshSimplicial3D M = UnitCube3D();
    for (int l = 0; l < levels; l++) M.uniformrefinement();
    M.automatic_dirichlet_flags();
    WhitneyResult& r = g_whitney;
    r = WhitneyResult();
    r.ne = M.count_simplices(1);
    r.edges.resize((size_t)r.ne * 2);
    r.dirichlet.resize(r.ne);
    for (int e = 0; e < r.ne; e++) {
        r.edges[2 * e] = M.get_subsimplex(1, 0, e, 0);
        r.edges[2 * e + 1] = M.get_subsimplex(1, 0, e, 1);
        r.dirichlet[e] = M.get_flag(1, e) == SimplexFlag::SimplexFlagDirichlet ? 1 : 0;
    }
    const MatrixCSR incl(FEECWhitneyInclusionMatrix(M, 3, 1, 1));
    const MatrixCSR diff(FEECBrokenDiffMatrix(M, 3, 1, 1));
    const MatrixCSR mass2(FEECBrokenMassMatrix(M, 3, 2, 0));
    const MatrixCSR mass1(FEECBrokenMassMatrix(M, 3, 1, 1));
    keep_csr(Conjugation(mass2, diff & incl), r.kA, r.kC, r.kV);
    keep_csr(Conjugation(mass1, incl), r.mA, r.mC, r.mV);
    return r.ne;
}

// Degree-r Lagrange stiffness in FEEC's Sullivan basis on UnitSquare2D_simple
// refined `levels` times with automatic Dirichlet flags, composed as
// tests/sullivan2D/poissondirichlet2D.cpp:120-151 composes it:
// (D I)^T M_1 (D I) with I = FEECSullivanInclusionMatrix(M, 2, 0, r),
// D = FEECBrokenDiffMatrix(M, 2, 0, r), M_1 = FEECBrokenMassMatrix(M, 2, 1, r-1),
// sortentries, MatrixCSR.  Also the mesh's edges (vertex pairs) and the
// Dirichlet flags of vertices and edges.  Returns the number of unknowns.
struct SullivanResult {
    int nv = 0, ne = 0, nt = 0, n = 0;
    std::vector<int> edges, vflag, eflag, triangles, tri_edges;
    std::vector<int> A, C;
    std::vector<double> V;
} g_sullivan;

int feecref_sullivan2d(int levels, int r)
{
    MeshSimplicial2D M = UnitSquare2D_simple();
    for (int l = 0; l < levels; l++) M.uniformrefinement();
    M.automatic_dirichlet_flags();
    SullivanResult& res = g_sullivan;
    res = SullivanResult();
    res.nv = M.count_simplices(0);
    res.ne = M.count_simplices(1);
    res.nt = M.count_simplices(2);
    for (int e = 0; e < res.ne; e++) {
        res.edges.push_back(M.get_subsimplex(1, 0, e, 0));
        res.edges.push_back(M.get_subsimplex(1, 0, e, 1));
        res.eflag.push_back(M.get_flag(1, e) == SimplexFlag::SimplexFlagDirichlet ? 1 : 0);
    }
    for (int v = 0; v < res.nv; v++)
        res.vflag.push_back(M.get_flag(0, v) == SimplexFlag::SimplexFlagDirichlet ? 1 : 0);
    for (int t = 0; t < res.nt; t++)
        for (int i = 0; i < 3; i++) {
            res.triangles.push_back(M.get_subsimplex(2, 0, t, i));
            res.tri_edges.push_back(M.get_subsimplex(2, 1, t, i));
        }
    SparseMatrix vector_massmatrix = FEECBrokenMassMatrix(M, 2, 1, r - 1);
    SparseMatrix diffmatrix = FEECBrokenDiffMatrix(M, 2, 0, r);
    SparseMatrix incmatrix = FEECSullivanInclusionMatrix(M, 2, 0, r);
    auto opr = diffmatrix & incmatrix;
    auto opl = opr.getTranspose();
    auto stiffness = opl & (vector_massmatrix & opr);
    stiffness.sortentries();
    const MatrixCSR csr(stiffness);
    keep_csr(csr, res.A, res.C, res.V);
    res.n = csr.getdimout();
    return res.n;
}

// sizes: vertices, edges, triangles, unknowns, entries
void feecref_sullivan2d_sizes(int* out)
{
    const SullivanResult& r = g_sullivan;
    out[0] = r.nv;
    out[1] = r.ne;
    out[2] = r.nt;
    out[3] = r.n;
    out[4] = r.A.empty() ? 0 : r.A.back();
}

void feecref_sullivan2d_copy(int* edges, int* vflag, int* eflag, int* triangles, int* tri_edges,
                             int* A, int* C, double* V)
{
    const SullivanResult& r = g_sullivan;
    std::memcpy(edges, r.edges.data(), r.edges.size() * sizeof(int));
    std::memcpy(vflag, r.vflag.data(), r.vflag.size() * sizeof(int));
    std::memcpy(eflag, r.eflag.data(), r.eflag.size() * sizeof(int));
    std::memcpy(triangles, r.triangles.data(), r.triangles.size() * sizeof(int));
    std::memcpy(tri_edges, r.tri_edges.data(), r.tri_edges.size() * sizeof(int));
    std::memcpy(A, r.A.data(), r.A.size() * sizeof(int));
    std::memcpy(C, r.C.data(), r.C.size() * sizeof(int));
    std::memcpy(V, r.V.data(), r.V.size() * sizeof(double));
}

// sizes: edges, entries of K, entries of M
void feecref_whitney1_sizes(int* out)
{
    out[0] = g_whitney.ne;
    out[1] = g_whitney.kA.empty() ? 0 : g_whitney.kA.back();
    out[2] = g_whitney.mA.empty() ? 0 : g_whitney.mA.back();
}

void feecref_whitney1_copy(int* edges, int* dirichlet, int* kA, int* kC, double* kV, int* mA,
                           int* mC, double* mV)
{
    const WhitneyResult& r = g_whitney;
    std::memcpy(edges, r.edges.data(), r.edges.size() * sizeof(int));
    std::memcpy(dirichlet, r.dirichlet.data(), r.dirichlet.size() * sizeof(int));
    std::memcpy(kA, r.kA.data(), r.kA.size() * sizeof(int));
    std::memcpy(kC, r.kC.data(), r.kC.size() * sizeof(int));
    std::memcpy(kV, r.kV.data(), r.kV.size() * sizeof(double));
    std::memcpy(mA, r.mA.data(), r.mA.size() * sizeof(int));
    std::memcpy(mC, r.mC.data(), r.mC.size() * sizeof(int));
    std::memcpy(mV, r.mV.data(), r.mV.size() * sizeof(double));
}

}  // extern "C"
