// GPU conjugate gradient with the reference LinearSolver contract, over the
// C ABI of libamie_gpu.so.
//
// Reference interface it stands beside:
//   LinearSolver::solve(const Vector &x0, Preconditionner *precond,
//                       double eps = default_solver_precision (1e-10),
//                       int maxit = -1, bool verbose = false)
//     solvers/solver.h:46-52, polynomial/variable.h:14
//   ConjugateGradient::solve                 solvers/conjugategradient.cpp:78-318
//   Assembly::getMatrix() / getForces()      solvers/assembly.h:337-346
//   CoordinateIndexedSparseMatrix fields     sparse/sparse_matrix.h:132-135
//
// Argument meaning follows the reference:
//   x0       the start vector; a shorter x0 fills the leading entries of x
//            (conjugategradient.cpp:104-113);
//   eps      absolute stop sqrt(r . D^-1 r) <= max(1e-12, eps), D = diag(K),
//            the reference's Jacobi-preconditioned residual (:101, :218, :305);
//   maxit    -1 = the number of unknowns: the reference sets Maxit = n/2
//            (:102) but tests it only between restarts, and one inner run
//            may take up to n iterations (:214), so it can exceed n/2;
//   a right-hand side with max |f| < eps^2 returns true at once (:81-85).
// Error behaviour: the reference prints to std::cerr and calls exit(0) on a
// NaN (:160-165); here the call prints the library's message and returns
// false.  rowstart / colstart (space-time rows) must be 0.
//
// One engine holds one GPU context and the system of its last solve on the
// device: a later solve whose matrix has the same pattern (Assembly::
// make_final keeps it while the mesh does, solvers/assembly.cpp:425-537)
// uploads only the new values and right-hand side.
//
// AssemblyT needs getMatrix() (fields stride, array, column_index, row_size:
// the reference CoordinateIndexedSparseMatrix) and getForces() (a
// std::valarray<double>).  The matrix is the final system of make_final(),
// Dirichlet rows included.  Header only; link with -lamie_gpu.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <valarray>
#include <vector>

#include "amie_gpu.h"

namespace amie {

class GpuConjugateGradientEngine {
public:
    explicit GpuConjugateGradientEngine(int device = 0)
    {
        if (amie_gpu_create(&ctx_, device) != AMIE_OK) {
            std::cerr << "GPU CG: " << amie_gpu_last_error() << std::endl;
            ctx_ = nullptr;
        }
    }
    ~GpuConjugateGradientEngine()
    {
        if (batch_) amie_gpu_batch_destroy(batch_);
        if (ctx_) amie_gpu_destroy(ctx_);
    }
    GpuConjugateGradientEngine(const GpuConjugateGradientEngine &) = delete;
    GpuConjugateGradientEngine &operator=(const GpuConjugateGradientEngine &) = delete;

    bool ready() const { return ctx_ != nullptr; }

    int precision = AMIE_MIXED_FIX48;   // AMIE_FP64 for the plain FP64 CG
    int precond = AMIE_JACOBI;          // AMIE_BLOCK_JACOBI for Inverse2x2Diagonal
    int64_t iterations = 0;             // of the last solve
    double relres = 0.;                 // ||b - K x|| / ||b|| of the last solve (FP64)
    int status = AMIE_OK;               // of the last solve
    bool reused = false;                // the last solve kept the device system (same pattern)
    int64_t uploads = 0;                // device systems built from scratch so far

    // Solves K x = b (K in the reference block-CSR layout) starting from x.
    template <class MatrixT>
    bool solve(const MatrixT &K, const std::valarray<double> &b, std::valarray<double> &x, double eps, int64_t maxit,
               bool verbose)
    {
        if (!ctx_) return false;
        const int stride = (int)K.stride;
        const int64_t nrows = (int64_t)K.row_size.size();
        if ((size_t)(nrows * stride) != b.size() || x.size() != b.size()) {
            std::cerr << "GPU CG: matrix has " << nrows * stride << " unknowns, forces " << b.size() << std::endl;
            return false;
        }
        reused = batch_ && same_pattern(K);
        if (reused) {
            status = amie_gpu_batch_update_bcsr(batch_, &K.array[0], &b[0]);
            if (status != AMIE_OK) return fail("amie_gpu_batch_update_bcsr");
        } else {
            if (batch_) amie_gpu_batch_destroy(batch_);
            batch_ = nullptr;
            status = amie_gpu_batch_from_bcsr(ctx_, stride, nrows, &K.row_size[0], &K.column_index[0], &K.array[0],
                                              &b[0], &batch_);
            if (status != AMIE_OK) return fail("amie_gpu_batch_from_bcsr");
            stride_ = stride;
            row_size_.assign(std::begin(K.row_size), std::end(K.row_size));
            column_index_.assign(std::begin(K.column_index), std::end(K.column_index));
            uploads++;
        }
        status = amie_gpu_batch_set_solution(batch_, 0, &x[0]);
        if (status != AMIE_OK) return fail("amie_gpu_batch_set_solution");
        amie_solve_opts o;
        amie_gpu_default_opts(&o);
        o.precision = precision;
        o.precond = precond;
        o.use_x0 = 1;
        o.maxit = maxit;
        o.ref_eps = eps;
        amie_solve_stats st;
        status = amie_gpu_batch_solve(batch_, &o, &st);
        if (status < 0) return fail("amie_gpu_batch_solve");
        iterations = st.iterations_max;
        relres = st.relres_max;
        if (amie_gpu_batch_get_solution(batch_, 0, &x[0]) != AMIE_OK) return fail("amie_gpu_batch_get_solution");
        if (verbose)
            std::cerr << "\n GPU CG " << b.size() << (status == AMIE_OK ? " converged after " : " did not converge after ")
                      << iterations << " iterations. Relative residual : " << relres << std::endl;
        return status == AMIE_OK;
    }

private:
    template <class MatrixT> bool same_pattern(const MatrixT &K) const
    {
        return (int)K.stride == stride_ && K.row_size.size() == row_size_.size() &&
               K.column_index.size() == column_index_.size() &&
               std::equal(row_size_.begin(), row_size_.end(), std::begin(K.row_size)) &&
               std::equal(column_index_.begin(), column_index_.end(), std::begin(K.column_index));
    }
    bool fail(const char *what)
    {
        std::cerr << "GPU CG: " << what << ": " << amie_gpu_last_error() << std::endl;
        if (batch_) amie_gpu_batch_destroy(batch_);
        batch_ = nullptr;
        return false;
    }
    amie_gpu_ctx *ctx_ = nullptr;
    amie_gpu_batch *batch_ = nullptr;   // the device system of the last solve
    int stride_ = 0;
    std::vector<uint32_t> row_size_, column_index_;
};

// LinearSolver::solve over an engine: the reference's handling of x0, eps,
// maxit, homogeneous right-hand sides and space-time rows, then the GPU
// solve of the assembly's final system into x (the solver's x member).
template <class AssemblyT>
bool linear_solver_solve(GpuConjugateGradientEngine &engine, AssemblyT *assembly, std::valarray<double> &x,
                         unsigned long rowstart, unsigned long colstart, const std::valarray<double> &x0, double eps,
                         int maxit, bool verbose)
{
    const std::valarray<double> &f = assembly->getForces();
    if (rowstart || colstart) {
        std::cerr << "GPU CG: space-time rows (rowstart/colstart) are not supported" << std::endl;
        return false;
    }
    if (x.size() != f.size()) x.resize(f.size(), 0.);
    if (f.size() == 0 || std::abs(f).max() < eps * eps) {
        if (verbose) std::cerr << "\n GPU CG " << x.size() << " homogeneous." << std::endl;
        return true;
    }
    if (x0.size() == f.size()) {
        x = x0;
    } else {
        for (size_t i = 0; i < std::min(f.size(), x0.size()); i++) x[i] = x0[i];
    }
    const int64_t mit = maxit != -1 ? maxit : (int64_t)f.size();
    return engine.solve(assembly->getMatrix(), f, x, eps, mit, verbose);
}

// The drop-in as a standalone class with the reference LinearSolver's
// public members (x, assembly, rowstart, colstart) for callers that do not
// derive from the reference's LinearSolver (tests with a stand-in
// Assembly).  PrecondT is the caller's preconditioner type; the GPU applies
// Jacobi unless engine.precond says block-Jacobi.
template <class AssemblyT, class PrecondT = void>
class GpuConjugateGradient {
public:
    typedef std::valarray<double> Vector;
    explicit GpuConjugateGradient(AssemblyT *a, int device = 0)
        : x(0., a->getForces().size()), assembly(a), engine(device)
    {
    }

    Vector x;
    AssemblyT *assembly;
    unsigned long rowstart = 0, colstart = 0;
    GpuConjugateGradientEngine engine;

    bool solve(const Vector &x0, PrecondT *precond = nullptr, const double eps = 1e-10, const int maxit = -1,
               bool verbose = false)
    {
        (void)precond;
        return linear_solver_solve(engine, assembly, x, rowstart, colstart, x0, eps, maxit, verbose);
    }
};

} // namespace amie
