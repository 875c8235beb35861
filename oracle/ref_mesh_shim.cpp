// ref_mesh_shim.cpp -- extern "C" entry points into the reference's own mesh
// refinement and P1 assembly, compiled from the reference sources by
// oracle/Makefile into oracle/_ref/libfeecref.so.
//
// TEST INFRASTRUCTURE ONLY: tests/golden/make_golden.py and tests/ use it to
// pin the build's reference-numbered mesh generator (vertex numbering,
// simplices, CSR structure and values).  The shim adds nothing to the
// algorithms: it builds the reference's unit square / unit cube, refines it,
// sets the Dirichlet flags and assembles the P1 stiffness matrices the
// reference's tests use.
#include <cstring>
#include <vector>

#include "base/include.hpp"
#include "dense/densematrix.hpp"
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

}  // extern "C"
