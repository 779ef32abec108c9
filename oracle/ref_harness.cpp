// Golden-vector harness for the compiled reference AMIE library (oracle/_ref).
//
// Test infrastructure only.  Built by oracle/Makefile.ref against the
// reference sources where they lie and run in the development container; the
// small .npy fixtures it writes under tests/golden/ref/ are committed and pin
// the CPU restatement in oracle/amie_oracle.cpp.  Everything below drives the
// reference's own public API:
//   Tensor::cauchyGreen                    utilities/tensor.cpp:1845-1931
//   DelaunayTriangle::getElementaryMatrix  mesher/delaunay.cpp:2408-2535
//   DelaunayTetrahedron::getElementaryMatrix mesher/delaunay_3d.cpp:2452
//   Assembly::add / setPointAlong / make_final / setBoundaryConditions
//                                          solvers/assembly.cpp:106-123,425-1223,1590
//   assign(r, A*x-b), assign(q, A*x)       sparse/sparse_matrix.cpp:462-547
//   InverseDiagonal, Inverse2x2Diagonal    solvers/inversediagonal.cpp:48-133
//   parallel_inner_product                 utilities/matrixops.h:246-262
//   ConjugateGradient::solve               solvers/conjugategradient.cpp:69-318
//
// Usage: ref_harness <outdir>

#include "elements/elements.h"
#include "features/features.h"
#include "features/inclusion.h"
#include "features/sample.h"
#include "features/sample3d.h"
#include "features/boundarycondition.h"
#include "utilities/writer/triangle_writer.h"
#include "utilities/placement.h"
#include "utilities/writer/voxel_writer.h"
#include "mesher/delaunay.h"
#include "mesher/delaunay_3d.h"
#include "physics/fracturecriteria/fracturecriterion.h"
#include "polynomial/vm_function_extra.h"
#include "physics/stiffness.h"
#include "physics/stiffness_and_fracture.h"
#include "physics/fracturecriteria/mazars.h"
#include "physics/damagemodels/fiberbasedisotropiclineardamage.h"
#include "physics/damagemodels/isotropiclineardamage.h"
#include "solvers/assembly.h"
#include "solvers/conjugategradient.h"
#include "solvers/inversediagonal.h"
#include "sparse/sparse_matrix.h"
#include "utilities/matrixops.h"
#include "utilities/tensor.h"

#include <omp.h>
#include <algorithm>
#include <functional>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

using namespace Amie;

static std::string g_out;

// ---- minimal .npy writer (format version 1.0, little endian) -------------
static void write_npy(const std::string &name, const char *descr, size_t itemsize,
                      const std::vector<size_t> &shape, const void *data)
{
    std::string shp = "(";
    size_t count = 1;
    for (size_t i = 0; i < shape.size(); i++) {
        shp += std::to_string(shape[i]);
        shp += (shape.size() == 1) ? "," : (i + 1 < shape.size() ? ", " : "");
        count *= shape[i];
    }
    shp += ")";
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + shp + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr += std::string(pad, ' ');
    hdr += "\n";
    std::string path = g_out + "/" + name + ".npy";
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(1); }
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    fwrite(magic, 1, 8, f);
    uint16_t hl = (uint16_t)hdr.size();
    fwrite(&hl, 2, 1, f);
    fwrite(hdr.data(), 1, hdr.size(), f);
    fwrite(data, itemsize, count, f);
    fclose(f);
}
static void npy_f64(const std::string &n, const std::vector<double> &v, std::vector<size_t> shape = {})
{
    if (shape.empty()) shape = {v.size()};
    write_npy(n, "<f8", 8, shape, v.data());
}
static void npy_u32(const std::string &n, const std::vector<uint32_t> &v)
{
    write_npy(n, "<u4", 4, {v.size()}, v.data());
}
static void npy_i64(const std::string &n, const std::vector<int64_t> &v, std::vector<size_t> shape = {})
{
    if (shape.empty()) shape = {v.size()};
    write_npy(n, "<i8", 8, shape, v.data());
}
static std::vector<double> vec(const Vector &v) { return std::vector<double>(&v[0], &v[0] + v.size()); }
template <class T> static std::vector<uint32_t> uvec(const std::valarray<T> &v)
{
    std::vector<uint32_t> r(v.size());
    for (size_t i = 0; i < v.size(); i++) r[i] = (uint32_t)v[i];
    return r;
}
static std::vector<double> mat(const Matrix &m)
{
    std::vector<double> r(m.numRows() * m.numCols());
    for (size_t i = 0; i < m.numRows(); i++)
        for (size_t j = 0; j < m.numCols(); j++) r[i * m.numCols() + j] = m[i][j];
    return r;
}

// Dense element matrix from the reference's block cache: dof = local_node*ndof + axis.
static std::vector<double> dense_elem(std::valarray<std::valarray<Matrix>> &K, size_t nn, size_t nd)
{
    std::vector<double> d(nn * nd * nn * nd);
    for (size_t i = 0; i < nn; i++)
        for (size_t j = 0; j < nn; j++)
            for (size_t a = 0; a < nd; a++)
                for (size_t b = 0; b < nd; b++) d[(i * nd + a) * (nn * nd) + j * nd + b] = K[i][j][a][b];
    return d;
}

// ---------------------------------------------------------------- C matrices
static void golden_cauchy_green()
{
    npy_f64("C_ps_E1e10_nu0.2", mat(Tensor::cauchyGreen(1e10, 0.2, SPACE_TWO_DIMENSIONAL, PLANE_STRESS, YOUNG_POISSON)), {3, 3});
    npy_f64("C_ps_E12e9_nu0.3", mat(Tensor::cauchyGreen(12e9, 0.3, SPACE_TWO_DIMENSIONAL, PLANE_STRESS, YOUNG_POISSON)), {3, 3});
    npy_f64("C_ps_E60e9_nu0.2", mat(Tensor::cauchyGreen(60e9, 0.2, SPACE_TWO_DIMENSIONAL, PLANE_STRESS, YOUNG_POISSON)), {3, 3});
    npy_f64("C_3d_E30e9_nu0.2", mat(Tensor::cauchyGreen(30e9, 0.2, SPACE_THREE_DIMENSIONAL, PLANE_STRESS, YOUNG_POISSON)), {6, 6});
}

// ----------------------------------------------------------- element matrices
static void golden_p1(const std::string &tag, double x0, double y0, double x1, double y1, double x2, double y2, double E, double nu)
{
    TriElement father(LINEAR);
    Point *p[3] = {new Point(x0, y0), new Point(x1, y1), new Point(x2, y2)};
    for (int i = 0; i < 3; i++) p[i]->getId() = i;
    DelaunayTriangle *t = new DelaunayTriangle(nullptr, nullptr, p[0], p[1], p[2], nullptr);
    t->refresh(&father);
    t->setBehaviour(nullptr, new Stiffness(E, nu));
    VirtualMachine vm;
    auto &K = t->getElementaryMatrix(&vm);
    std::vector<int64_t> order(3);
    for (int i = 0; i < 3; i++) order[i] = t->getBoundingPoint(i).getId();
    npy_f64("Ke_" + tag, dense_elem(K, 3, 2), {6, 6});
    npy_i64("Ke_" + tag + "_order", order);
    npy_f64("Ke_" + tag + "_xy", {x0, y0, x1, y1, x2, y2}, {3, 2});
}

static void golden_p2(const std::string &tag, double x0, double y0, double x1, double y1, double x2, double y2, double E, double nu)
{
    // P2 bounding points follow the shape-function layout of
    // TriElement(QUADRATIC) (elements/elements.cpp:1002-1061):
    // [v0, m(v0,v1), v1, m(v1,v2), v2, m(v2,v0)], with v0..v2 already CCW.
    TriElement father(QUADRATIC);
    Point *v[3] = {new Point(x0, y0), new Point(x1, y1), new Point(x2, y2)};
    for (int i = 0; i < 3; i++) v[i]->getId() = 2 * i;
    DelaunayTriangle *t = new DelaunayTriangle(nullptr, nullptr, v[0], v[1], v[2], nullptr);
    Point *a = &t->getBoundingPoint(0), *b = &t->getBoundingPoint(1), *c = &t->getBoundingPoint(2);
    Point *m01 = new Point(0.5 * (a->getX() + b->getX()), 0.5 * (a->getY() + b->getY()));
    Point *m12 = new Point(0.5 * (b->getX() + c->getX()), 0.5 * (b->getY() + c->getY()));
    Point *m20 = new Point(0.5 * (c->getX() + a->getX()), 0.5 * (c->getY() + a->getY()));
    m01->getId() = 1; m12->getId() = 3; m20->getId() = 5;
    std::valarray<Point *> pts(6);
    pts[0] = a; pts[1] = m01; pts[2] = b; pts[3] = m12; pts[4] = c; pts[5] = m20;
    t->setBoundingPoints(pts);
    t->refresh(&father);
    t->setBehaviour(nullptr, new Stiffness(E, nu));
    VirtualMachine vm;
    auto &K = t->getElementaryMatrix(&vm);
    std::vector<double> xy;
    for (int i = 0; i < 6; i++) { xy.push_back(t->getBoundingPoint(i).getX()); xy.push_back(t->getBoundingPoint(i).getY()); }
    npy_f64("Ke_" + tag, dense_elem(K, 6, 2), {12, 12});
    npy_f64("Ke_" + tag + "_xy", xy, {6, 2});
}

static void golden_tet(const std::string &tag, const double *c, double E, double nu)
{
    TetrahedralElement father(LINEAR);
    Point *p[4];
    for (int i = 0; i < 4; i++) { p[i] = new Point(c[3 * i], c[3 * i + 1], c[3 * i + 2]); p[i]->getId() = i; }
    DelaunayTetrahedron *t = new DelaunayTetrahedron(nullptr, nullptr, p[0], p[1], p[2], p[3], nullptr);
    t->refresh(&father);
    t->setBehaviour(nullptr, new Stiffness(E, nu, SPACE_THREE_DIMENSIONAL));
    VirtualMachine vm;
    auto &K = t->getElementaryMatrix(&vm);
    std::vector<int64_t> order(4);
    for (int i = 0; i < 4; i++) order[i] = t->getBoundingPoint(i).getId();
    npy_f64("Ke_" + tag, dense_elem(K, 4, 3), {12, 12});
    npy_i64("Ke_" + tag + "_order", order);
    npy_f64("Ke_" + tag + "_xyz", std::vector<double>(c, c + 12), {4, 3});
}

// ------------------------------------------------- structured meshes + systems
// P1: node (i,j) -> id i*(ny+1)+j, x = i*lx/nx, y = j*ly/ny; two triangles per
// cell as mesher/structuredmesh.cpp:26-31 (no jitter).  Phase of a triangle is
// given by `phase_of` (0 = E0/nu0, 1 = E1/nu1).
struct P1Case {
    int nx, ny;
    double lx, ly;
    double E0, nu0, E1, nu1;
};

static void golden_system_p1(const std::string &tag, const P1Case &pc, int nssor, bool solve)
{
    TriElement father(LINEAR);
    std::vector<Point *> pts;
    for (int i = 0; i <= pc.nx; i++)
        for (int j = 0; j <= pc.ny; j++) {
            pts.push_back(new Point(pc.lx * i / pc.nx, pc.ly * j / pc.ny));
            pts.back()->getId() = (int)pts.size() - 1;
        }
    auto P = [&](int i, int j) { return pts[i * (pc.ny + 1) + j]; };
    Assembly K;
    std::vector<DelaunayTriangle *> tris;
    std::vector<int64_t> conn, phase;
    for (int i = 0; i < pc.nx; i++)
        for (int j = 0; j < pc.ny; j++) {
            Point *c0[3] = {P(i, j), P(i, j + 1), P(i + 1, j + 1)};
            Point *c1[3] = {P(i, j), P(i + 1, j), P(i + 1, j + 1)};
            for (int h = 0; h < 2; h++) {
                Point **c = h ? c1 : c0;
                DelaunayTriangle *t = new DelaunayTriangle(nullptr, nullptr, c[0], c[1], c[2], nullptr);
                t->refresh(&father);
                int ph = ((i * 7 + j * 3 + h) % 5 == 0) ? 1 : 0;
                if (pc.E1 == pc.E0) ph = 0;
                t->setBehaviour(nullptr, ph ? new Stiffness(pc.E1, pc.nu1) : new Stiffness(pc.E0, pc.nu0));
                tris.push_back(t);
                for (int k = 0; k < 3; k++) conn.push_back(t->getBoundingPoint(k).getId());
                phase.push_back(ph);
                K.add(t, 1.);
            }
        }
    K.setMaxDof(pts.size());
    // BCs of the uniaxial configuration: bottom edge u_y = 0, corner (0,0)
    // u_x = 0, top edge u_y = 1e-3.
    std::vector<int64_t> bc_dof;
    std::vector<double> bc_val;
    for (int i = 0; i <= pc.nx; i++) {
        K.setPointAlong(ETA, 0., P(i, 0)->getId());
        bc_dof.push_back(2 * P(i, 0)->getId() + 1); bc_val.push_back(0.);
        K.setPointAlong(ETA, 1e-3, P(i, pc.ny)->getId());
        bc_dof.push_back(2 * P(i, pc.ny)->getId() + 1); bc_val.push_back(1e-3);
    }
    K.setPointAlong(XI, 0., P(0, 0)->getId());
    bc_dof.push_back(0); bc_val.push_back(0.);

    // Raw (pre-BC) matrix: a second assembly without multipliers.
    Assembly Kraw;
    for (auto t : tris) Kraw.add(t, 1.);
    Kraw.setMaxDof(pts.size());
    Kraw.make_final();
    for (auto t : tris) t->behaviourUpdated = true;

    K.make_final();
    CoordinateIndexedSparseMatrix &A = K.getMatrix();
    std::vector<double> xy;
    for (auto p : pts) { xy.push_back(p->getX()); xy.push_back(p->getY()); }
    npy_f64(tag + "_xy", xy, {pts.size(), 2});
    npy_i64(tag + "_conn", conn, {tris.size(), 3});
    npy_i64(tag + "_phase", phase);
    npy_f64(tag + "_mat", {pc.E0, pc.nu0, pc.E1, pc.nu1});
    npy_i64(tag + "_bc_dof", bc_dof);
    npy_f64(tag + "_bc_val", bc_val);
    npy_u32(tag + "_row_size", uvec(A.row_size));
    npy_u32(tag + "_column_index", uvec(A.column_index));
    npy_f64(tag + "_array", vec(A.array));
    npy_f64(tag + "_raw_array", vec(Kraw.getMatrix().array));
    npy_f64(tag + "_forces", vec(K.getForces()));

    // SpMV A*x - b and A*x with a seeded x.
    size_t n = K.getForces().size();
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1., 1.);
    Vector x(n), r(n), q(n);
    for (size_t i = 0; i < n; i++) x[i] = U(rng);
    assign(r, A * x - K.getForces());
    assign(q, A * x);
    npy_f64(tag + "_spmv_x", vec(x));
    npy_f64(tag + "_spmv_Axmb", vec(r));
    npy_f64(tag + "_spmv_Ax", vec(q));
    Vector t(n);
    InverseDiagonal D(A);
    D.precondition(x, t);
    npy_f64(tag + "_invdiag_x", vec(t));
    Inverse2x2Diagonal D2(A);
    D2.precondition(x, t);
    npy_f64(tag + "_inv2x2_x", vec(t));
    npy_f64(tag + "_dot_x_Ax", {parallel_inner_product(&x[0], &q[0], (int)n)});

    if (solve) {
        ConjugateGradient cg(&K);
        cg.nssor = nssor;
        Vector x0(0., n);
        bool ok = cg.solve(x0, nullptr, 1e-10, -1, false);
        npy_f64(tag + "_cg_x", vec(cg.x));
        npy_i64(tag + "_cg_info", {(int64_t)ok, (int64_t)cg.nit, (int64_t)nssor});
    }
}

// Tet4 structured cube mesh, 6-tet split of filters/voxelfilter.cpp:352-405,
// node id i*(n+1)^2 + j*(n+1) + k.
static void golden_system_tet(const std::string &tag, int n, double E, double nu)
{
    TetrahedralElement father(LINEAR);
    std::vector<Point *> pts;
    int m = n + 1;
    for (int i = 0; i <= n; i++)
        for (int j = 0; j <= n; j++)
            for (int k = 0; k <= n; k++) {
                pts.push_back(new Point((double)i / n, (double)j / n, (double)k / n));
                pts.back()->getId() = (int)pts.size() - 1;
            }
    auto P = [&](int i, int j, int k) { return pts[(size_t)i * m * m + j * m + k]; };
    Assembly K;
    std::vector<DelaunayTetrahedron *> tets;
    std::vector<int64_t> conn;
    static const int split[6][4] = {{0, 1, 2, 4}, {3, 5, 7, 6}, {2, 3, 4, 6}, {1, 2, 3, 4}, {1, 3, 4, 5}, {3, 4, 6, 5}};
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                // corners c0..c7 of the voxel (filters/voxelfilter.cpp:352-359)
                Point *c[8] = {P(i, j + 1, k), P(i + 1, j + 1, k), P(i, j + 1, k + 1), P(i + 1, j + 1, k + 1),
                               P(i, j, k), P(i + 1, j, k), P(i, j, k + 1), P(i + 1, j, k + 1)};
                for (int s = 0; s < 6; s++) {
                    DelaunayTetrahedron *t = new DelaunayTetrahedron(nullptr, nullptr, c[split[s][0]], c[split[s][1]], c[split[s][2]], c[split[s][3]], nullptr);
                    t->refresh(&father);
                    t->setBehaviour(nullptr, new Stiffness(E, nu, SPACE_THREE_DIMENSIONAL));
                    tets.push_back(t);
                    for (int q = 0; q < 4; q++) conn.push_back(t->getBoundingPoint(q).getId());
                    K.add(t, 1.);
                }
            }
    K.setMaxDof(pts.size());
    std::vector<int64_t> bc_dof;
    std::vector<double> bc_val;
    for (int i = 0; i <= n; i++)
        for (int j = 0; j <= n; j++) {
            for (int a = 0; a < 3; a++) {
                Variable v = a == 0 ? XI : (a == 1 ? ETA : ZETA);
                K.setPointAlong(v, 0., P(i, j, 0)->getId());
                bc_dof.push_back(3 * P(i, j, 0)->getId() + a); bc_val.push_back(0.);
            }
            K.setPointAlong(ZETA, 2e-5, P(i, j, n)->getId());
            bc_dof.push_back(3 * P(i, j, n)->getId() + 2); bc_val.push_back(2e-5);
        }
    K.make_final();
    CoordinateIndexedSparseMatrix &A = K.getMatrix();
    std::vector<double> xyz;
    for (auto p : pts) { xyz.push_back(p->getX()); xyz.push_back(p->getY()); xyz.push_back(p->getZ()); }
    npy_f64(tag + "_xyz", xyz, {pts.size(), 3});
    npy_i64(tag + "_conn", conn, {tets.size(), 4});
    npy_f64(tag + "_mat", {E, nu});
    npy_i64(tag + "_bc_dof", bc_dof);
    npy_f64(tag + "_bc_val", bc_val);
    npy_u32(tag + "_row_size", uvec(A.row_size));
    npy_u32(tag + "_column_index", uvec(A.column_index));
    npy_f64(tag + "_array", vec(A.array));
    npy_f64(tag + "_forces", vec(K.getForces()));
    size_t nd = K.getForces().size();
    std::mt19937_64 rng(777);
    std::uniform_real_distribution<double> U(-1., 1.);
    Vector x(nd), r(nd);
    for (size_t i = 0; i < nd; i++) x[i] = U(rng);
    assign(r, A * x - K.getForces());
    npy_f64(tag + "_spmv_x", vec(x));
    npy_f64(tag + "_spmv_Axmb", vec(r));
    ConjugateGradient cg(&K);
    cg.nssor = 32;
    Vector x0(0., nd);
    bool ok = cg.solve(x0, nullptr, 1e-10, -1, false);
    npy_f64(tag + "_cg_x", vec(cg.x));
    npy_i64(tag + "_cg_info", {(int64_t)ok, (int64_t)cg.nit, 32});
}

namespace Amie {
// physics/fracturecriteria/fracturecriterion.cpp:43-61 (not declared in a header)
Function getSmoothingKernelFunction(SmoothingFunctionType type, Function &rrn);
}

// The coefficients of Mesh::generateCache are a protected member and
// DelaunayTree is final: they are read through a pointer to member formed in
// an explicit instantiation (where access is not checked).
typedef Mesh<DelaunayTriangle, DelaunayTreeItem> TriMesh;
struct CoefsTag {
    typedef std::vector<std::vector<std::vector<double>>> TriMesh::*type;
    friend type member(CoefsTag);
};
template <typename Tag, typename Tag::type M> struct Expose {
    friend typename Tag::type member(Tag) { return M; }
};
template struct Expose<CoefsTag, &TriMesh::coefs>;

// Non-local neighbourhoods and weights: Mesh::generateCache
// (mesher/mesh.h:546-640) with the locus, radius and kernel that
// FractureCriterion::initialiseCache uses (fracturecriterion.cpp:171-190), on
// a Delaunay mesh of seeded points in the unit right triangle.  No element has
// a behaviour, so with a non-null source none is skipped and every weight
// comes from the smoothing kernel.
static void golden_smoothing(const std::string &tag, int npts, double radius, SmoothingFunctionType type)
{
    TriElement father(LINEAR);
    Point *c0 = new Point(0., 0.), *c1 = new Point(1., 0.), *c2 = new Point(0., 1.);
    DelaunayTree tree(c0, c1, c2);
    std::mt19937_64 rng(2024);
    std::uniform_real_distribution<double> U(0.02, 0.98);
    std::vector<Point *> pts = {c0, c1, c2};
    while ((int)pts.size() < npts + 3) {
        const double x = U(rng), y = U(rng);
        if (x + y > 0.97) continue;
        Point *p = new Point(x, y);
        pts.push_back(p);
        tree.insert(p, 0.);
    }
    std::vector<DelaunayTriangle *> tris = tree.getTriangles();
    std::map<const Point *, int64_t> pid;
    std::map<int, int64_t> pos;   // tree index -> position in tris
    std::vector<double> xy;
    std::vector<int64_t> conn;
    for (size_t e = 0; e < tris.size(); e++) {
        tris[e]->refresh(&father);
        pos[tris[e]->index] = (int64_t)e;
        for (int k = 0; k < 3; k++) {
            const Point *p = &tris[e]->getBoundingPoint(k);
            auto it = pid.find(p);
            if (it == pid.end()) {
                it = pid.emplace(p, (int64_t)pid.size()).first;
                xy.push_back(p->getX());
                xy.push_back(p->getY());
            }
            conn.push_back(it->second);
        }
    }
    Circle source(1., 0., 0.);
    std::vector<int64_t> ptr = {0}, idx;
    std::vector<double> w;
    for (size_t e = 0; e < tris.size(); e++) {
        DelaunayTriangle *t = tris[e];
        const double R = std::max(radius, t->getRadius() * 1.1);
        Function x = Function("x") - t->getCenter().getX();
        Function y = Function("y") - t->getCenter().getY();
        Function rr = x * x + y * y;
        Function rrn = rr / (R * R);
        Function smooth = getSmoothingKernelFunction(type, rrn);
        Circle locus(R * 2.5 + t->getRadius() * 1.5, t->getCenter());
        const unsigned id = tree.generateCache(&locus, &source, smooth);
        std::vector<int> &cache = tree.getCache(id);
        std::vector<std::vector<double>> &cw = (static_cast<TriMesh &>(tree).*member(CoefsTag()))[id];
        for (size_t k = 0; k < cache.size(); k++) {
            idx.push_back(pos.at(cache[k]));
            w.push_back(cw[k].at(0));
        }
        ptr.push_back((int64_t)idx.size());
        cache.clear();
        cw.clear();
    }
    npy_f64(tag + "_xy", xy, {xy.size() / 2, 2});
    npy_i64(tag + "_conn", conn, {tris.size(), 3});
    npy_i64(tag + "_ptr", ptr);
    npy_i64(tag + "_idx", idx);
    npy_f64(tag + "_w", w);
    npy_f64(tag + "_radius", {radius});
}

// FeatureTree's own mesh pipeline up to renumbering: the state machine of
// features/features.cpp:5767-5920 (sample, generateElements, stitch -- which
// for QUADRATIC adds the mid-edge nodes, mesher/delaunay.cpp:260-355 --,
// element behaviours, renumber, features/features.cpp:603-729).  FeatureTree's
// `state` and its StateType are protected: the member pointer is formed in an
// explicit instantiation (access not checked there) and the enumerator is
// passed by value (RENUMBERED = 3, features/features.h:58-72).
auto feature_tree_state();
template <auto M> struct ExposeAuto {
    friend auto feature_tree_state() { return M; }
};
template struct ExposeAuto<&FeatureTree::state>;

// setStateTo's parameter type deduced from its member pointer, so that the
// protected enumeration need not be named.
template <typename S, typename E> static void call_set_state(S &st, void (S::*f)(E, bool), int v)
{
    (st.*f)(static_cast<E>(v), false);
}

template <typename T> static void renumber_tree(T &F)
{
    auto &st = F.*feature_tree_state();
    call_set_state(st, &std::remove_reference_t<decltype(st)>::setStateTo, 3 /* RENUMBERED */);
}

// Elements in the mesh's iteration order (the order renumber() walks), each
// with the ids renumber() gave its bounding points, and the coordinates by id.
template <typename MeshT> static void dump_mesh(const std::string &tag, MeshT *m, int dim)
{
    std::vector<int64_t> conn;
    std::map<int, std::vector<double>> xyz;
    size_t nper = 0;
    for (auto i = m->begin(); i != m->end(); i++) {
        nper = i->getBoundingPoints().size();
        for (size_t j = 0; j < nper; j++) {
            const Point &p = i->getBoundingPoint(j);
            conn.push_back(p.getId());
            std::vector<double> c = {p.getX(), p.getY()};
            if (dim == 3) c.push_back(p.getZ());
            xyz[p.getId()] = c;
        }
    }
    std::vector<double> coords;
    int expect = 0;
    for (auto &kv : xyz) {
        if (kv.first != expect++) { fprintf(stderr, "%s: ids not contiguous\n", tag.c_str()); exit(1); }
        coords.insert(coords.end(), kv.second.begin(), kv.second.end());
    }
    npy_i64(tag + "_conn", conn, {conn.size() / nper, nper});
    npy_f64(tag + "_coords", coords, {xyz.size(), (size_t)dim});
}

static void golden_ingest_2d(const std::string &tag, double sampling, Order order, bool inclusion)
{
    RectangularFeature sample(nullptr, 1., 1., 0.5, 0.5);
    sample.setBehaviour(new Stiffness(1e10, 0.2));
    FeatureTree F(&sample);
    Inclusion *inc = nullptr;
    if (inclusion) {
        inc = new Inclusion(&sample, 0.21, 0.45, 0.55);
        inc->setBehaviour(new Stiffness(6e10, 0.2));
        F.addFeature(&sample, inc);
    }
    F.setSamplingNumber(sampling);
    F.setOrder(order);
    F.setProjectionOnBoundaries(false);
    renumber_tree(F);
    dump_mesh(tag, F.get2DMesh(), 2);
}

static void golden_ingest_3d(const std::string &tag, double sampling)
{
    Sample3D sample(nullptr, 1., 1., 1., 0.5, 0.5, 0.5);
    sample.setBehaviour(new Stiffness(1e10, 0.2, SPACE_THREE_DIMENSIONAL));
    FeatureTree F(&sample);
    F.setSamplingNumber(sampling);
    F.setOrder(LINEAR);
    renumber_tree(F);
    dump_mesh(tag, F.get3DMesh(), 3);
}

// VoxelWriter (utilities/writer/voxel_writer.cpp:52-272, 433-442, 548-571)
// after a FeatureTree solve of a unit cube clamped on its left face and
// sheared on its right (u_y = 1e-3), a non-uniform stress field: the six
// VWFT_STRESS files it writes, and its inputs -- the tetrahedra in mesh
// order and each one's stress (constant over a Tet4; taken at its centre).
static void golden_voxels(const std::string &tag, double sampling, int nvox)
{
    Sample3D sample(nullptr, 1., 1., 1., 0.5, 0.5, 0.5);
    sample.setBehaviour(new Stiffness(1e10, 0.2, SPACE_THREE_DIMENSIONAL));
    FeatureTree F(&sample);
    F.setSamplingNumber(sampling);
    F.setOrder(LINEAR);
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_XI, LEFT));
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_ETA, LEFT));
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_ZETA, LEFT));
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(SET_ALONG_ETA, RIGHT, 1e-3));
    F.step();
    VoxelWriter vw(g_out + "/" + tag, nvox);
    vw.getField(&F, VWFT_STRESS);
    vw.write();
    dump_mesh(tag, F.get3DMesh(), 3);
    std::vector<double> sig;
    VirtualMachine vm;
    auto *m = F.get3DMesh();
    for (auto i = m->begin(); i != m->end(); i++) {
        Vector t(0., 6);
        i->getState().getField(REAL_STRESS_FIELD, i->getCenter(), t, false, &vm);
        for (int c = 0; c < 6; c++) sig.push_back(t[c]);
    }
    npy_f64(tag + "_stress", sig, {sig.size() / 6, 6});
}

// placement2D (utilities/placement.cpp:87-162) of circular inclusions in a
// rectangle: radii drawn from a seeded list (largest first), placed with the
// given gap, tries budget and seed; which were placed, where, and with what
// radius after the gap scaling is undone.
static void golden_placement(const std::string &tag, int n, double rmin, double rmax, double min_dist, int tries,
                             size_t seed)
{
    RectangularFeature sample(nullptr, 1., 0.8, 0.5, 0.4);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> ur(rmin, rmax);
    std::vector<double> radii(n);
    for (auto &r : radii) r = ur(rng);
    std::sort(radii.begin(), radii.end(), std::greater<double>());
    std::vector<Feature *> incs;
    for (double r : radii) incs.push_back(new Inclusion(r, 0., 0.));
    std::vector<Feature *> placed = placement2D(sample.getPrimitive(), incs, min_dist, 0, tries, M_PI,
                                                std::vector<Geometry *>(), seed);
    std::vector<int64_t> idx;
    std::vector<double> out;
    for (Feature *f : placed) {
        idx.push_back(std::find(incs.begin(), incs.end(), f) - incs.begin());
        out.push_back(f->getCenter().getX());
        out.push_back(f->getCenter().getY());
        out.push_back(f->getRadius());
    }
    npy_f64(tag + "_radii", radii);
    npy_i64(tag + "_placed", idx);
    npy_f64(tag + "_out", out, {idx.size(), 3});
    npy_f64(tag + "_params", {min_dist, (double)tries, (double)seed});
}

// TriangleWriter (utilities/writer/triangle_writer.cpp:31-86, 407-445,
// 578-608, 1024-1075, 1461-1475), BinaryTriangleWriter (:447-478,
// 1477-1491) and MultiTriangleWriter after a FeatureTree solve of a unit box
// (u_x = 0 left, u_y = 0 bottom, u_x = 1e-3 right): the file it writes, and
// the inputs it read -- mesh, DOF displacements, and the displacement,
// strain and stress the element states give at each element's three
// vertices (interpolated there: a fixed DOF can read back as 1e-20).
static void golden_writer(const std::string &tag, double sampling)
{
    RectangularFeature sample(nullptr, 1., 1., 0.5, 0.5);
    sample.setBehaviour(new Stiffness(1e10, 0.2));
    FeatureTree F(&sample);
    F.setSamplingNumber(sampling);
    F.setOrder(LINEAR);
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_XI, LEFT));
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_ETA, BOTTOM));
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(SET_ALONG_XI, RIGHT, 1e-3));
    F.step();
    TriangleWriter trg(g_out + "/" + tag + "_trg.txt", &F, 1);
    trg.getField(TOTAL_STRAIN_FIELD);
    trg.getField(REAL_STRESS_FIELD);
    trg.write();
    BinaryTriangleWriter btrg(g_out + "/" + tag + "_btrg.txt", &F, 1);
    btrg.getField(TOTAL_STRAIN_FIELD);
    btrg.getField(REAL_STRESS_FIELD);
    btrg.write();
    // MultiTriangleWriter (:73-86, 538-573, 1493-1520): an index file and one
    // TriangleWriter file per append; two appends of the same fields
    MultiTriangleWriter mtrg(g_out + "/" + tag + "_multi_head.txt", g_out + "/" + tag + "_multi", &F, 1);
    mtrg.getField(TOTAL_STRAIN_FIELD);
    mtrg.getField(REAL_STRESS_FIELD);
    mtrg.append();
    mtrg.append();
    dump_mesh(tag, F.get2DMesh(), 2);
    npy_f64(tag + "_u", vec(F.getDisplacements()));
    std::vector<double> eps, sig, uv;
    VirtualMachine vm;
    auto *m = F.get2DMesh();
    for (auto i = m->begin(); i != m->end(); i++)
        for (int k = 0; k < 3; k++) {
            Vector e(3), t(3), d(2);
            i->getState().getField(TOTAL_STRAIN_FIELD, i->getBoundingPoint(k), e, false, &vm);
            i->getState().getField(REAL_STRESS_FIELD, i->getBoundingPoint(k), t, false, &vm);
            i->getState().getField(DISPLACEMENT_FIELD, i->getBoundingPoint(k), d, false, &vm, 0);
            for (int c = 0; c < 3; c++) { eps.push_back(e[c]); sig.push_back(t[c]); }
            uv.push_back(d[0]);
            uv.push_back(d[1]);
        }
    npy_f64(tag + "_uvertex", uv, {uv.size() / 6, 3, 2});
    npy_f64(tag + "_strain", eps, {eps.size() / 9, 3, 3});
    npy_f64(tag + "_stress", sig, {sig.size() / 9, 3, 3});
}

// The reference's event-driven damage through FeatureTree::step
// (features/features.cpp:6150-6373, stepElements :4875-5290): a unit square
// of StiffnessAndFracture(E, nu, NonLocalMazars(threshold, E, nu, Gf,
// radius)) with the default FiberBasedIsotropicLinearDamage (+0.1 per event,
// fractured at 0.6, residual stiffness 1e-4; physics/damagemodels/
// fiberbasedisotropiclineardamage.cpp:14-100) around a stiffer elastic
// inclusion, u_x = 0 left, u_y = 0 bottom, u_x = load[k] right, one
// FeatureTree::step per load.  After each step: every element's damage and
// fracture score (mesh order), the displacements and the number of damage
// model step() calls (passes x damaging elements); once, the order the
// first pass visits the elements in (element positions; every pass visits
// them in the same order) and, per call of that pass, the element's score,
// its neighbourhood's maximum score and its smoothed principal strains and
// stresses.
struct LoggedFiberDamage : public FiberBasedIsotropicLinearDamage {
    static std::vector<int64_t> *calls;
    static std::map<const IntegrableEntity *, int64_t> *position;
    // the first pass in full: per call the element's score, the maximum
    // score of its neighbourhood and its smoothed principal strains and
    // stresses, as the criterion sees them
    static std::vector<double> *first;
    static size_t first_calls;
    static FILE *trace;  // AMIE_EVENT_TRACE: one line per call (debugging only, not a fixture)
    LoggedFiberDamage(double f, double c) : FiberBasedIsotropicLinearDamage(f, c) { }
    virtual void step(ElementState &s, double maxscore)
    {
        FractureCriterion *fc = s.getParent()->getBehaviour()->getFractureCriterion();
        const bool record = fc && ((first && first->size() < first_calls * 6) || trace);
        const double before = getState()[0];
        const double nbr = record ? fc->getMaxScoreInNeighbourhood(s) : 0.;
        std::pair<Vector, Vector> sf;
        if (record) sf = fc->getSmoothedFields(PRINCIPAL_REAL_STRESS_FIELD, PRINCIPAL_TOTAL_STRAIN_FIELD, s, 0.);
        FiberBasedIsotropicLinearDamage::step(s, maxscore);
        const int64_t pos = position && position->count(s.getParent()) ? (*position)[s.getParent()] : -1;
        if (calls) calls->push_back(pos);
        if (record && first && first->size() < first_calls * 6)
            first->insert(first->end(), {fc->getScoreAtState(), nbr, sf.second[0], sf.second[1], sf.first[0], sf.first[1]});
        if (record && trace) {
            fprintf(trace, "%lld %.17g %.17g %.17g %.17g", (long long)pos, fc->getScoreAtState(), nbr, before, getState()[0]);
            for (size_t q = 0; q < sf.first.size(); q++) fprintf(trace, " %.17g", sf.first[q]);
            for (size_t q = 0; q < sf.second.size(); q++) fprintf(trace, " %.17g", sf.second[q]);
            fprintf(trace, "\n");
        }
    }
    virtual DamageModel *getCopy() const
    {
        LoggedFiberDamage *ret = new LoggedFiberDamage(fibreFraction, getThresholdDamageDensity());
        ret->copyEssentialParameters(this);
        return ret;
    }
};
std::vector<int64_t> *LoggedFiberDamage::calls = nullptr;
std::map<const IntegrableEntity *, int64_t> *LoggedFiberDamage::position = nullptr;
std::vector<double> *LoggedFiberDamage::first = nullptr;
size_t LoggedFiberDamage::first_calls = 0;
FILE *LoggedFiberDamage::trace = nullptr;

static void golden_events(const std::string &tag, double sampling, const std::vector<double> &loads, double E,
                          double nu, double threshold, double Gf, double radius, double E_inc)
{
    RectangularFeature sample(nullptr, 1., 1., 0.5, 0.5);
    sample.setBehaviour(new StiffnessAndFracture(E, nu, new NonLocalMazars(threshold, E, nu, Gf, radius, PLANE_STRESS),
                                                 new LoggedFiberDamage(0.1, 0.6)));
    FeatureTree F(&sample);
    Inclusion *inc = new Inclusion(&sample, 0.18, 0.5, 0.5);
    inc->setBehaviour(new Stiffness(E_inc, nu));
    F.addFeature(&sample, inc);
    F.setSamplingNumber(sampling);
    F.setOrder(LINEAR);
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_XI, LEFT));
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_ETA, BOTTOM));
    BoundingBoxDefinedBoundaryCondition *right = new BoundingBoxDefinedBoundaryCondition(SET_ALONG_XI, RIGHT, 0.);
    F.addBoundaryCondition(right);
    renumber_tree(F);
    auto *m = F.get2DMesh();
    std::map<const IntegrableEntity *, int64_t> pos;
    std::vector<int64_t> phase;
    for (auto i = m->begin(); i != m->end(); i++) {
        pos[&*i] = i.getPosition();
        phase.push_back(i->getBehaviour()->getFractureCriterion() ? 0 : 1);
    }
    std::vector<int64_t> calls, ncalls;
    std::vector<double> dmg, score, u, first;
    int64_t ncrit = 0;
    for (int64_t v : phase) ncrit += v == 0;
    LoggedFiberDamage::position = &pos;
    LoggedFiberDamage::calls = &calls;
    LoggedFiberDamage::first = &first;
    LoggedFiberDamage::first_calls = (size_t)ncrit;
    if (getenv("AMIE_EVENT_TRACE")) LoggedFiberDamage::trace = fopen(getenv("AMIE_EVENT_TRACE"), "w");
    for (double v : loads) {
        right->setData(v);
        const size_t before = calls.size();
        F.step();
        ncalls.push_back((int64_t)(calls.size() - before));
        for (auto i = m->begin(); i != m->end(); i++) {
            DamageModel *dm = i->getBehaviour()->getDamageModel();
            FractureCriterion *fc = i->getBehaviour()->getFractureCriterion();
            dmg.push_back(dm ? dm->getState()[0] : 0.);
            score.push_back(fc ? fc->getScoreAtState() : 0.);
        }
        Vector x = F.getDisplacements();
        u.insert(u.end(), std::begin(x), std::end(x));
    }
    LoggedFiberDamage::calls = nullptr;
    LoggedFiberDamage::position = nullptr;
    LoggedFiberDamage::first = nullptr;
    if (LoggedFiberDamage::trace) fclose(LoggedFiberDamage::trace);
    LoggedFiberDamage::trace = nullptr;
    const size_t ne = phase.size(), ns = loads.size();
    dump_mesh(tag, m, 2);
    npy_i64(tag + "_phase", phase);
    npy_f64(tag + "_loads", loads);
    npy_f64(tag + "_params", {E, nu, threshold, Gf, radius, E_inc});
    npy_f64(tag + "_damage", dmg, {ns, ne});
    npy_f64(tag + "_score", score, {ns, ne});
    npy_f64(tag + "_u", u, {ns, u.size() / ns});
    npy_i64(tag + "_ncalls", ncalls);
    npy_i64(tag + "_order", std::vector<int64_t>(calls.begin(), calls.begin() + ncrit));
    npy_f64(tag + "_first", first, {first.size() / 6, 6});
}

// The same loop with IsotropicLinearDamage as the damage model (its Newton
// branch, physics/damagemodels/isotropiclineardamage.cpp:64-92 and
// postProcess :150-162).  The model is final, so the passes are observed at
// the criterion: a NonLocalMazars whose grade() (one call per element and
// pass, FractureCriterion::step, fracturecriterion.cpp:534-560) records the
// element's position, its damage before the pass and its score.
struct CountingMazars : public NonLocalMazars {
    struct Call { int64_t pos; double state, score; };
    static std::vector<Call> *calls;
    static std::map<const IntegrableEntity *, int64_t> *position;
    CountingMazars(double thr, double E_, double nu_, double Gf_, double cs, double ce, double r, planeType p)
        : NonLocalMazars(thr, E_, nu_, Gf_, cs, ce, r, p) { }
    CountingMazars(double thr, double E_, double nu_, double Gf_, double r, planeType p)
        : NonLocalMazars(thr, E_, nu_, Gf_, r, p) { }
    virtual double grade(ElementState &s)
    {
        const double g = NonLocalMazars::grade(s);
        if (calls && position)
            calls->push_back({position->count(s.getParent()) ? position->at(s.getParent()) : -1,
                              s.getParent()->getBehaviour()->getDamageModel()->getState()[0], g});
        return g;
    }
    virtual FractureCriterion *getCopy() const
    {
        CountingMazars *ret = new CountingMazars(threshold, E, nu, Gf, cstress, cstrain, getMaterialCharacteristicRadius(), pt);
        ret->copyEssentialParameters(this);
        return ret;
    }
};
std::vector<CountingMazars::Call> *CountingMazars::calls = nullptr;
std::map<const IntegrableEntity *, int64_t> *CountingMazars::position = nullptr;

static void golden_events_newton(const std::string &tag, double sampling, const std::vector<double> &loads, double E,
                                 double nu, double threshold, double Gf, double radius, double E_inc)
{
    RectangularFeature sample(nullptr, 1., 1., 0.5, 0.5);
    sample.setBehaviour(new StiffnessAndFracture(E, nu, new CountingMazars(threshold, E, nu, Gf, radius, PLANE_STRESS),
                                                 new IsotropicLinearDamage()));
    FeatureTree F(&sample);
    Inclusion *inc = new Inclusion(&sample, 0.18, 0.5, 0.5);
    inc->setBehaviour(new Stiffness(E_inc, nu));
    F.addFeature(&sample, inc);
    F.setSamplingNumber(sampling);
    F.setOrder(LINEAR);
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_XI, LEFT));
    F.addBoundaryCondition(new BoundingBoxDefinedBoundaryCondition(FIX_ALONG_ETA, BOTTOM));
    BoundingBoxDefinedBoundaryCondition *right = new BoundingBoxDefinedBoundaryCondition(SET_ALONG_XI, RIGHT, 0.);
    F.addBoundaryCondition(right);
    renumber_tree(F);
    auto *m = F.get2DMesh();
    std::vector<int64_t> phase, ncalls;
    std::map<const IntegrableEntity *, int64_t> pos;
    for (auto i = m->begin(); i != m->end(); i++) {
        phase.push_back(i->getBehaviour()->getFractureCriterion() ? 0 : 1);
        pos[&*i] = i.getPosition();
    }
    std::vector<double> dmg, score, u;
    std::vector<CountingMazars::Call> calls;
    CountingMazars::calls = &calls;
    CountingMazars::position = &pos;
    for (double v : loads) {
        right->setData(v);
        const size_t before = calls.size();
        F.step();
        ncalls.push_back(int64_t(calls.size() - before));
        for (auto i = m->begin(); i != m->end(); i++) {
            DamageModel *dm = i->getBehaviour()->getDamageModel();
            FractureCriterion *fc = i->getBehaviour()->getFractureCriterion();
            dmg.push_back(dm ? dm->getState()[0] : 0.);
            score.push_back(fc ? fc->getScoreAtState() : 0.);
        }
        Vector x = F.getDisplacements();
        u.insert(u.end(), std::begin(x), std::end(x));
    }
    CountingMazars::calls = nullptr;
    CountingMazars::position = nullptr;
    // selected passes: the damage before pass k and k + 1 and the scores of
    // pass k, for every element (0 outside the criterion phase); a pass is
    // one grade() call per criterion element
    const size_t ne = phase.size(), ns = loads.size();
    size_t ncrit = 0;
    for (int64_t p : phase) ncrit += p == 0;
    const size_t npass = calls.size() / ncrit;
    std::vector<int64_t> sel, sel_load;
    std::vector<double> st0, sc0, st1;
    size_t first = 0;
    for (size_t k = 0; k < ns; k++) {
        const size_t np = size_t(ncalls[k]) / ncrit;
        for (size_t j : {size_t(0), size_t(1), size_t(2), np / 4, np / 2, np - 2, np - 1}) {
            const size_t g = first + j;
            if (j >= np || g + 1 >= npass || (!sel.empty() && sel.back() >= int64_t(g))) continue;
            sel.push_back(int64_t(g));
            sel_load.push_back(int64_t(k));
            for (size_t w = 0; w < 2; w++) {
                std::vector<double> a(ne, 0.), b(ne, 0.);
                for (size_t c = (g + w) * ncrit; c < (g + w + 1) * ncrit; c++) {
                    a[size_t(calls[c].pos)] = calls[c].state;
                    b[size_t(calls[c].pos)] = calls[c].score;
                }
                if (w == 0) { st0.insert(st0.end(), a.begin(), a.end()); sc0.insert(sc0.end(), b.begin(), b.end()); }
                else st1.insert(st1.end(), a.begin(), a.end());
            }
        }
        first += np;
    }
    dump_mesh(tag, m, 2);
    npy_i64(tag + "_phase", phase);
    npy_f64(tag + "_loads", loads);
    npy_f64(tag + "_params", {E, nu, threshold, Gf, radius, E_inc});
    npy_f64(tag + "_damage", dmg, {ns, ne});
    npy_f64(tag + "_score", score, {ns, ne});
    npy_f64(tag + "_u", u, {ns, u.size() / ns});
    npy_i64(tag + "_pass", sel);
    npy_i64(tag + "_pass_load", sel_load);
    npy_f64(tag + "_pass_state", st0, {sel.size(), ne});
    npy_f64(tag + "_pass_score", sc0, {sel.size(), ne});
    npy_f64(tag + "_pass_next", st1, {sel.size(), ne});
    npy_i64(tag + "_ncalls", ncalls);
}

// The reference's own ConjugateGradient::solve (solvers/conjugategradient.cpp:
// 69-318, default InverseDiagonal preconditioner, nssor 32 as
// Assembly::cgsolve sets it, solvers/assembly.h:262 and assembly.cpp:1842)
// on a system read from a file, with every OpenMP thread the environment
// gives; the CPU baseline of the benchmark.  File (native byte order):
// int32 stride, int64 nrows, uint32 row_size[nrows], int64 nnzb,
// uint32 column_index[nnzb], double array[nnzb*stride*colLen],
// double rhs[nrows*stride], double eps (absolute, on sqrt(r.z)), int64 maxit.
// Writes the solution to <file>.x and prints "iterations seconds converged
// threads" on one line; seconds cover the solve only.
static int run_cg(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); return 2; }
    int32_t stride = 0;
    int64_t nrows = 0, nnzb = 0, maxit = 0;
    bool ok = fread(&stride, 4, 1, f) == 1 && fread(&nrows, 8, 1, f) == 1;
    std::valarray<unsigned int> rs(ok ? (size_t)nrows : 0);
    ok = ok && fread(&rs[0], 4, nrows, f) == (size_t)nrows && fread(&nnzb, 8, 1, f) == 1;
    std::valarray<unsigned int> ci(ok ? (size_t)nnzb : 0);
    ok = ok && fread(&ci[0], 4, nnzb, f) == (size_t)nnzb;
    const size_t cl = stride + stride % 2, nval = (size_t)nnzb * stride * cl, n = (size_t)nrows * stride;
    std::vector<double> vals(ok ? nval : 0), rhs(ok ? n : 0);
    double eps = 0.;
    ok = ok && fread(vals.data(), 8, nval, f) == nval && fread(rhs.data(), 8, n, f) == n && fread(&eps, 8, 1, f) == 1 &&
         fread(&maxit, 8, 1, f) == 1;
    fclose(f);
    if (!ok) { fprintf(stderr, "short file %s\n", path); return 2; }
    Assembly K;
    K.coordinateIndexedMatrix = new CoordinateIndexedSparseMatrix(rs, ci, stride);
    for (size_t q = 0; q < nval; q++) K.coordinateIndexedMatrix->array[q] = vals[q];
    K.externalForces.resize(n);
    for (size_t q = 0; q < n; q++) K.externalForces[q] = rhs[q];
    ConjugateGradient cg(&K);
    cg.nssor = 32;
    Vector x0(0., n);
    const double t0 = omp_get_wtime();
    const bool conv = cg.solve(x0, nullptr, eps, (int)maxit, false);
    const double dt = omp_get_wtime() - t0;
    std::string out = std::string(path) + ".x";
    FILE *g = fopen(out.c_str(), "wb");
    if (!g || fwrite(&cg.x[0], 8, n, g) != n) { fprintf(stderr, "cannot write %s\n", out.c_str()); return 2; }
    fclose(g);
    printf("%lld %.9f %d %d\n", (long long)cg.nit, dt, (int)conv, omp_get_max_threads());
    return 0;
}

// --cg-rate FILE K: the work of K iterations of ConjugateGradient::solve's
// inner loop (solvers/conjugategradient.cpp:212-253) on the system in FILE
// (layout of --cg), through the reference's own operations: InverseDiagonal
// precondition, parallel_inner_product_restricted, the p update, the block
// SpMV assign(q, A*p), and the Kahan-compensated x and r updates.  Prints
// "K seconds threads".  (The solver's own maxit is only checked between its
// restarts, so it cannot bound a solve of a large system to K iterations.)
static int run_cg_rate(const char *path, int iters)
{
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); return 2; }
    int32_t stride = 0;
    int64_t nrows = 0, nnzb = 0, maxit = 0;
    bool ok = fread(&stride, 4, 1, f) == 1 && fread(&nrows, 8, 1, f) == 1;
    std::valarray<unsigned int> rs(ok ? (size_t)nrows : 0);
    ok = ok && fread(&rs[0], 4, nrows, f) == (size_t)nrows && fread(&nnzb, 8, 1, f) == 1;
    std::valarray<unsigned int> ci(ok ? (size_t)nnzb : 0);
    ok = ok && fread(&ci[0], 4, nnzb, f) == (size_t)nnzb;
    const size_t cl = stride + stride % 2, nval = (size_t)nnzb * stride * cl, n = (size_t)nrows * stride;
    std::vector<double> vals(ok ? nval : 0), rhs(ok ? n : 0);
    double eps = 0.;
    ok = ok && fread(vals.data(), 8, nval, f) == nval && fread(rhs.data(), 8, n, f) == n && fread(&eps, 8, 1, f) == 1 &&
         fread(&maxit, 8, 1, f) == 1;
    fclose(f);
    if (!ok) { fprintf(stderr, "short file %s\n", path); return 2; }
    CoordinateIndexedSparseMatrix A(rs, ci, stride);
    for (size_t q = 0; q < nval; q++) A.array[q] = vals[q];
    InverseDiagonal P(A);
    Vector r(n), z(n), p(0., n), q(n), x(0., n), rc(0., n), xc(0., n);
    for (size_t i = 0; i < n; i++) r[i] = rhs[i];
    double last_rho = 1.;
    const int vsize = (int)n;
    const double t0 = omp_get_wtime();
    for (int it = 0; it < iters; it++) {
        P.precondition(r, z);
        const double rho = parallel_inner_product_restricted(&r[0], &z[0], vsize);
        const double beta = rho / last_rho;
#pragma omp parallel for
        for (int i = 0; i < vsize; i++) p[i] = p[i] * beta + z[i];
        assign(q, A * p, 0, 0);
        const double pq = parallel_inner_product_restricted(&q[0], &p[0], vsize);
        const double alpha = pq != 0. ? rho / pq : 0.;
#pragma omp parallel for
        for (int i = 0; i < vsize; i++) {
            const double yr = -q[i] * alpha - rc[i], yx = p[i] * alpha - xc[i];
            const double rt = r[i] + yr, xt = x[i] + yx;
            rc[i] = (rt - r[i]) - yr;
            xc[i] = (xt - x[i]) - yx;
            r[i] = rt;
            x[i] = xt;
        }
        last_rho = rho;
    }
    const double dt = omp_get_wtime() - t0;
    printf("%d %.9f %d\n", iters, dt, omp_get_max_threads());
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 3 && std::string(argv[1]) == "--cg") return run_cg(argv[2]);
    if (argc == 4 && std::string(argv[1]) == "--cg-rate") return run_cg_rate(argv[2], atoi(argv[3]));
    if (argc == 3 && std::string(argv[1]) == "--events") {
        g_out = argv[2];
        omp_set_num_threads(1);
        golden_events("events_p1", 10, {0.8e-4, 1.2e-4, 1.6e-4, 2.0e-4}, 1e10, 0.2, 1e-4, 100., 0.05, 4e10);
        // a smaller radius: elements fractured within a pass change which
        // later-visited elements are at their neighbourhood's maximum
        golden_events("events_r01", 10, {1.1e-4, 1.375e-4, 1.65e-4}, 1e10, 0.2, 1e-4, 30., 0.01, 4e10);
        golden_events_newton("events_newton", 10, {0.8e-4, 1.2e-4, 1.6e-4}, 1e10, 0.2, 1e-4, 100., 0.05, 4e10);
        printf("event-damage golden vectors written to %s\n", g_out.c_str());
        return 0;
    }
    if (argc < 2) { fprintf(stderr, "usage: %s outdir | --events outdir | --cg system.bin\n", argv[0]); return 2; }
    g_out = argv[1];
    omp_set_num_threads(1);
    golden_cauchy_green();
    golden_p1("p1_a", 0.1, 0.2, 0.9, 0.35, 0.3, 0.8, 1e10, 0.2);
    golden_p1("p1_cw", 0.0, 0.0, 0.0, 0.5, 0.25, 0.0, 12e9, 0.3); // needs CCW reorientation
    golden_p1("p1_unit", 0.0, 0.0, 0.0, 1.0 / 128, 1.0 / 128, 1.0 / 128, 1e10, 0.2);
    golden_p2("p2_a", 0.1, 0.2, 0.9, 0.35, 0.3, 0.8, 60e9, 0.2);
    golden_p2("p2_unit", 0.0, 1.0 / 400, 0.0, 0.0, 1.0 / 400, 0.0, 12e9, 0.3);
    const double t1[12] = {0.0, 0.0, 0.0, 1.0, 0.1, 0.0, 0.2, 0.9, 0.1, 0.3, 0.2, 1.1};
    const double t2[12] = {0.0, 0.0, 0.0, 0.2, 0.9, 0.1, 1.0, 0.1, 0.0, 0.3, 0.2, 1.1}; // other orientation
    golden_tet("tet_a", t1, 30e9, 0.2);
    golden_tet("tet_b", t2, 30e9, 0.2);
    golden_system_p1("sys_p1_4x3", {4, 3, 1.0, 1.0, 1e10, 0.2, 60e9, 0.2}, 32, true);
    golden_system_p1("sys_p1_8x8", {8, 8, 1.0, 1.0, 1e10, 0.2, 1e10, 0.2}, 32, true);
    golden_system_p1("sys_p1_8x8_nos", {8, 8, 1.0, 1.0, 1e10, 0.2, 1e10, 0.2}, 0, true);
    golden_system_p1("sys_p1_12x10_mix", {12, 10, 1.0, 1.0, 12e9, 0.3, 60e9, 0.2}, 32, true);
    golden_system_tet("sys_tet_2", 2, 30e9, 0.2);
    golden_system_tet("sys_tet_3", 3, 30e9, 0.2);
    golden_smoothing("smooth_quartic", 150, 0.05, QUARTIC_COMPACT);
    golden_smoothing("smooth_gauss", 150, 0.05, GAUSSIAN_NONCOMPACT);
    golden_ingest_2d("ingest_p1", 16, LINEAR, false);
    golden_ingest_2d("ingest_p2", 16, QUADRATIC, false);
    golden_ingest_2d("ingest_p2_inc", 24, QUADRATIC, true);
    golden_ingest_3d("ingest_tet", 4);
    golden_writer("writer_p1", 8);
    golden_voxels("voxels_tet", 4, 24);
    golden_placement("place_gap", 60, 0.01, 0.05, 0.005, 6400, 1);
    golden_placement("place_nogap", 200, 0.02, 0.08, 0., 3000, 5);
    printf("golden vectors written to %s\n", g_out.c_str());
    return 0;
}
