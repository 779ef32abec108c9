// The drop-in solver driven by the compiled reference itself: test
// infrastructure (built by oracle/Makefile.ref into oracle/_ref/, run by
// tests/test_gpu_parity.py on the GPU box).  It compiles the reference-side
// file xfem-amie_amd/host/gpuconjugategradient.h (the Amie::
// GpuConjugateGradient : LinearSolver that a maintainer adds to the
// reference tree) against the reference's own headers, builds systems with
// the reference's Assembly (solvers/assembly.cpp:106-123,425-1223,1590-1634)
// and solves each twice:
//   reference  Assembly::cgsolve()                 solvers/assembly.cpp:1829-1953
//   drop-in    Assembly::mgprepare() + mgsolve(&gcg, x0)   :1225-1228,1955-1969
// over three load steps per case (the same elements re-added and the top
// load raised, as FeatureTree::solve re-adds them every step), one
// GpuConjugateGradient object for all steps.  Checks: the displacements of
// every step against cgsolve's to 1e-6 (both stop on the reference's
// absolute sqrt(r . D^-1 r) <= 1e-10), one device system built for the
// three steps (the pattern does not change) and the later steps reusing it.
//
// Usage: ref_dropin PRECISION     (amie_gpu.h precision code)
// Exit status: 0 all checks pass, 1 a check failed, 2 the GPU is missing.

#include "elements/elements.h"
#include "mesher/delaunay.h"
#include "mesher/delaunay_3d.h"
#include "physics/stiffness.h"
#include "solvers/assembly.h"
#include "solvers/conjugategradient.h"

#include "gpuconjugategradient.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <vector>

using namespace Amie;

namespace {

double rel_diff(const Vector &a, const Vector &b)
{
    if (a.size() != b.size()) return 1e300;
    double d = 0., n = 0.;
    for (size_t i = 0; i < a.size(); i++) {
        d += (a[i] - b[i]) * (a[i] - b[i]);
        n += b[i] * b[i];
    }
    return std::sqrt(d / n);
}

// one load step on assembly K: elements added, boundary conditions set
using StepFn = std::function<void(Assembly &, double)>;

bool run_case(const char *name, size_t ndof_nodes, const StepFn &load_step, int precision)
{
    Assembly Kr, Kg;
    Kr.setMaxDof(ndof_nodes);
    Kg.setMaxDof(ndof_nodes);
    // made after the first mgprepare, as Assembly::cgsolve makes its solver
    // after make_final (the Solver constructor reads getForces(), which
    // would run make_final on an empty assembly)
    std::unique_ptr<GpuConjugateGradient> gcgp;
    bool ok = true;
    for (int s = 1; s <= 3; s++) {
        const double top = 1e-3 * s;
        load_step(Kr, top);
        if (!Kr.cgsolve()) {
            fprintf(stderr, "%s step %d: the reference CG did not converge\n", name, s);
            return false;
        }
        load_step(Kg, top);
        Kg.mgprepare();
        if (!gcgp) {
            gcgp.reset(new GpuConjugateGradient(&Kg));
            if (!gcgp->engine.ready()) exit(2);
            gcgp->engine.precision = precision;
        }
        GpuConjugateGradient &gcg = *gcgp;
        const bool conv = Kg.mgsolve(&gcg, Kg.getDisplacements());
        const double err = rel_diff(Kg.getDisplacements(), Kr.getDisplacements());
        printf("%s step %d: converged %d, iterations %lld, rel. diff vs cgsolve %.3e, device system %s (built %lld)\n",
               name, s, (int)conv, (long long)gcg.engine.iterations, err, gcg.engine.reused ? "reused" : "new",
               (long long)gcg.engine.uploads);
        ok &= conv && err <= 1e-6 && (s == 1 ? !gcg.engine.reused : gcg.engine.reused);
    }
    ok &= gcgp->engine.uploads == 1;
    return ok;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: ref_dropin PRECISION\n");
        return 2;
    }
    const int precision = atoi(argv[1]);

    // P1: unit square 24 x 20 (node (i,j) -> i*21 + j), two phases, bottom
    // u_y = 0, corner u_x = 0, top u_y = load
    const int nx = 24, ny = 20;
    TriElement tri_father(LINEAR);
    std::vector<Point *> pts;
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j <= ny; j++) {
            pts.push_back(new Point((double)i / nx, (double)j / ny));
            pts.back()->getId() = (int)pts.size() - 1;
        }
    auto P = [&](int i, int j) { return pts[i * (ny + 1) + j]; };
    std::vector<DelaunayTriangle *> tris;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            Point *c0[3] = {P(i, j), P(i, j + 1), P(i + 1, j + 1)};
            Point *c1[3] = {P(i, j), P(i + 1, j), P(i + 1, j + 1)};
            for (int h = 0; h < 2; h++) {
                Point **c = h ? c1 : c0;
                DelaunayTriangle *t = new DelaunayTriangle(nullptr, nullptr, c[0], c[1], c[2], nullptr);
                t->refresh(&tri_father);
                const bool inc = (i * 7 + j * 3 + h) % 5 == 0;
                t->setBehaviour(nullptr, inc ? new Stiffness(60e9, 0.2) : new Stiffness(12e9, 0.3));
                tris.push_back(t);
            }
        }
    // the element matrices are computed afresh for each assembly
    StepFn p1_step = [&](Assembly &K, double top) {
        for (auto t : tris) {
            t->behaviourUpdated = true;
            K.add(t, 1.);
        }
        for (int i = 0; i <= nx; i++) {
            K.setPointAlong(ETA, 0., P(i, 0)->getId());
            K.setPointAlong(ETA, top, P(i, ny)->getId());
        }
        K.setPointAlong(XI, 0., P(0, 0)->getId());
    };

    // Tet4: unit cube 5^3, 6-tet split, bottom clamped, top u_z = load / 50
    const int n = 5, m = n + 1;
    TetrahedralElement tet_father(LINEAR);
    std::vector<Point *> q;
    for (int i = 0; i <= n; i++)
        for (int j = 0; j <= n; j++)
            for (int k = 0; k <= n; k++) {
                q.push_back(new Point((double)i / n, (double)j / n, (double)k / n));
                q.back()->getId() = (int)q.size() - 1;
            }
    auto Q = [&](int i, int j, int k) { return q[(size_t)i * m * m + j * m + k]; };
    static const int split[6][4] = {{0, 1, 2, 4}, {3, 5, 7, 6}, {2, 3, 4, 6}, {1, 2, 3, 4}, {1, 3, 4, 5}, {3, 4, 6, 5}};
    std::vector<DelaunayTetrahedron *> tets;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                // corners c0..c7 of the voxel (filters/voxelfilter.cpp:352-359)
                Point *c[8] = {Q(i, j + 1, k), Q(i + 1, j + 1, k), Q(i, j + 1, k + 1), Q(i + 1, j + 1, k + 1),
                               Q(i, j, k), Q(i + 1, j, k), Q(i, j, k + 1), Q(i + 1, j, k + 1)};
                for (int h = 0; h < 6; h++) {
                    DelaunayTetrahedron *t = new DelaunayTetrahedron(nullptr, nullptr, c[split[h][0]], c[split[h][1]],
                                                                     c[split[h][2]], c[split[h][3]], nullptr);
                    t->refresh(&tet_father);
                    t->setBehaviour(nullptr, new Stiffness(30e9, 0.2, SPACE_THREE_DIMENSIONAL));
                    tets.push_back(t);
                }
            }
    StepFn tet_step = [&](Assembly &K, double top) {
        for (auto t : tets) {
            t->behaviourUpdated = true;
            K.add(t, 1.);
        }
        for (int i = 0; i <= n; i++)
            for (int j = 0; j <= n; j++) {
                for (Variable v : {XI, ETA, ZETA}) K.setPointAlong(v, 0., Q(i, j, 0)->getId());
                K.setPointAlong(ZETA, top / 50., Q(i, j, n)->getId());
            }
    };

    bool ok = run_case("P1 24x20", pts.size(), p1_step, precision);
    ok &= run_case("Tet4 5^3", q.size(), tet_step, precision);
    printf("%s\n", ok ? "ALL OK" : "FAILED");
    return ok ? 0 : 1;
}
