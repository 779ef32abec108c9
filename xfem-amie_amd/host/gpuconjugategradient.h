// solvers/gpuconjugategradient.h -- the file a maintainer adds to the
// reference tree (next to solvers/conjugategradient.h) to use the MI355X
// solver as a LinearSolver (solvers/solver.h:46-52), for instance through
// Assembly::mgsolve (solvers/assembly.cpp:1955-1969):
//
//   K->mgprepare();                             // K: Assembly*; make_final
//   Amie::GpuConjugateGradient gcg(K);          // after the first make_final:
//                                               // the Solver constructor reads
//                                               // K->getForces() (solvers/solver.cpp:16-18)
//   K->mgsolve(&gcg, K->getDisplacements());    // x0 = previous displacements
//
// Build with -I<repo>/include -I<repo>/xfem-amie_amd/host and link
// -lamie_gpu.  The solver object keeps one GPU context and the device
// system between solves: keep one per Assembly and reuse it across load
// and damage steps (a matrix with an unchanged pattern is updated in place
// on the device).  An Inverse2x2Diagonal preconditioner selects block
// Jacobi on the GPU, anything else Jacobi, the reference default
// (InverseDiagonal, solvers/conjugategradient.cpp:87-90); the passed
// preconditioner is never deleted.
#pragma once
#include "solver.h"
#include "inversediagonal.h"
#include "gpu_conjugate_gradient.hpp"

namespace Amie {

struct GpuConjugateGradient : public LinearSolver
{
    amie::GpuConjugateGradientEngine engine;
    explicit GpuConjugateGradient(Assembly * a, int device = 0) : LinearSolver(a), engine(device) { }
    virtual ~GpuConjugateGradient() { }

    virtual bool solve(const Vector & x0, Preconditionner * precond = nullptr,
                       const double eps = default_solver_precision, const int maxit = -1,
                       bool verbose = false)
    {
        engine.precond = dynamic_cast<Inverse2x2Diagonal *>(precond) ? AMIE_BLOCK_JACOBI : AMIE_JACOBI;
        return amie::linear_solver_solve(engine, assembly, x, rowstart, colstart, x0, eps, maxit, verbose);
    }
};

}
