// Drives amie::GpuConjugateGradient the way the reference's Assembly::mgsolve
// (solvers/assembly.cpp:1955-1969) drives a LinearSolver, on a system given as
// raw files in the reference block-CSR layout.  Used by the tests and as the
// compile check of the drop-in header.
//
//   dropin_check STRIDE ROW_SIZE.u32 COLUMN_INDEX.u32 ARRAY.f64 FORCES.f64 EPS PRECISION OUT.f64
//
// Exit status: 0 converged, 1 not converged, 2 error.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <valarray>

#include "gpu_conjugate_gradient.hpp"

namespace {

template <class T> bool read_file(const char *path, std::valarray<T> &out)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamsize n = f.tellg();
    if (n <= 0 || n % (std::streamsize)sizeof(T)) return false;
    out.resize((size_t)n / sizeof(T));
    f.seekg(0);
    return (bool)f.read(reinterpret_cast<char *>(&out[0]), n);
}

// The reference CoordinateIndexedSparseMatrix's public fields
// (sparse/sparse_matrix.h:132-135) and the two Assembly accessors the solver
// uses (solvers/assembly.h:337-346).
struct Matrix {
    size_t stride = 2;
    std::valarray<double> array;
    std::valarray<unsigned int> column_index, row_size;
};
struct Assembly {
    Matrix m;
    std::valarray<double> forces;
    Matrix &getMatrix() { return m; }
    std::valarray<double> &getForces() { return forces; }
};
struct Preconditionner;

} // namespace

int main(int argc, char **argv)
{
    if (argc != 9) {
        std::cerr << "usage: dropin_check STRIDE ROW_SIZE COLUMN_INDEX ARRAY FORCES EPS PRECISION OUT" << std::endl;
        return 2;
    }
    Assembly a;
    a.m.stride = (size_t)std::atoi(argv[1]);
    if (!read_file(argv[2], a.m.row_size) || !read_file(argv[3], a.m.column_index) || !read_file(argv[4], a.m.array) ||
        !read_file(argv[5], a.forces)) {
        std::cerr << "dropin_check: cannot read the input files" << std::endl;
        return 2;
    }
    amie::GpuConjugateGradient<Assembly, Preconditionner> cg(&a);
    if (!cg.engine.ready()) return 2;
    cg.engine.precision = std::atoi(argv[7]);
    const bool ok = cg.solve(std::valarray<double>(), nullptr, std::atof(argv[6]), -1, true);
    if (cg.engine.status < 0) return 2;
    std::ofstream o(argv[8], std::ios::binary);
    o.write(reinterpret_cast<const char *>(&cg.x[0]), (std::streamsize)(cg.x.size() * sizeof(double)));
    return ok ? 0 : 1;
}
