// CPU oracle: FP64 C++ restatement of the reference AMIE hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load this library, and only as the checker (or
// the timed CPU baseline) -- never as the product.  Every function names the
// reference file:line it restates.  Operation order follows the reference
// where that fixes the rounding (Kahan loops, fma calls, matrix products), so
// that with one OpenMP thread the results match the compiled reference's
// golden vectors in tests/golden/ref bit for bit or to the last ulp.
//
// Pinning (see tests/test_oracle_golden.py, tests/test_oracle_math.py):
//   or_cauchy_green      reference fixture examples/test/test_cauchy_green_base
//                        + compiled-reference golden vectors
//   or_orient            compiled-reference golden (local order)
//   or_element_matrix    compiled-reference golden; rigid-body null space;
//                        closed-form CST stiffness; patch tests
//   or_pattern           compiled-reference golden (bit-exact); brute force
//   or_assemble          compiled-reference golden; K * rigid modes = 0
//   or_apply_dirichlet   compiled-reference golden; closed-form elimination
//   or_spmv / or_dot     compiled-reference golden; dense numpy
//   or_inverse_*         compiled-reference golden; closed-form inverses
//   or_cg_reference      compiled-reference golden (iteration count, x)
//   or_pcg               scipy sparse direct solve; exact patch-test solution
//   or_element_strain    exact linear displacement fields
//   or_damage_update     closed form of the law; principal strains of
//                        numpy.linalg.eigvalsh; monotonicity
//   or_average_stress    homogeneous fields; Voigt/Reuss bounds
//   or_smoothed_strain   constant fields reproduced; kernel values at known
//                        distances; a lone element returns its own strain;
//                        phase and fracture exclusions; neighbourhood radius;
//                        the reference's own smoothed principal strains
//   or_gauss_fields      quadratic fields exact at the P2 Gauss points;
//                        uniaxial stress
//   or_mazars_score,     two FeatureTree runs of the compiled reference
//   or_fiber_pass        (damage, passes, scores, neighbourhood maxima);
//                        closed-form uniaxial score; a hand-made pass
//   or_newton_pass       twenty passes of a compiled-reference FeatureTree
//                        run with IsotropicLinearDamage replayed bit for bit
//                        from its damage and scores; its first load step
//                        (201 passes); a hand-made pass
//   or_principal         numpy.linalg.eigvalsh

#include "amie_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Shape {
    int nper, dim, b;
};
Shape shape_of(int etype)
{
    if (etype == 1) return {3, 2, 2};
    if (etype == 2) return {6, 2, 2};
    return {4, 3, 3};
}

// Shape-function derivatives at reference point (x, y[, z]).
// P1: elements/elements.cpp:976-1000 (N0 = eta, N1 = 1-xi-eta, N2 = xi).
// P2: elements/elements.cpp:1002-1061, evaluated in the order of the
//     reverse-polish derivative strings given there.
// Tet4: elements/elements.cpp:2491-2530 (N0 = zeta, N1 = 1-xi-eta-zeta,
//     N2 = xi, N3 = eta).
void dshape(int etype, double x, double y, double z, double *dxi, double *deta, double *dzeta)
{
    (void)z;
    if (etype == 1) {
        dxi[0] = 0.; deta[0] = 1.;
        dxi[1] = -1.; deta[1] = -1.;
        dxi[2] = 1.; deta[2] = 0.;
    } else if (etype == 2) {
        dxi[0] = 0.;                        deta[0] = 4. * y - 1.;
        dxi[1] = y * -4.;                   deta[1] = (4. - x * 4.) - y * 8.;
        dxi[2] = (y * 4. + x * 4.) - 3.;    deta[2] = (x * 4. + y * 4.) - 3.;
        dxi[3] = (4. - 8. * x) - y * 4.;    deta[3] = x * -4.;
        dxi[4] = 4. * x - 1.;               deta[4] = 0.;
        dxi[5] = 4. * y;                    deta[5] = 4. * x;
    } else {
        dxi[0] = 0.; deta[0] = 0.; dzeta[0] = 1.;
        dxi[1] = -1.; deta[1] = -1.; dzeta[1] = -1.;
        dxi[2] = 1.; deta[2] = 0.; dzeta[2] = 0.;
        dxi[3] = 0.; deta[3] = 1.; dzeta[3] = 0.;
    }
}

// utilities/matrixops.cpp:536-561
void invert2x2(double *s)
{
    if (std::abs(s[0]) < 1e-24) {
        double a = s[0], b = s[1], c = s[2], d = s[3];
        double dt = 1. / (a * d - b * c);
        s[0] = d * dt; s[1] = -b * dt; s[2] = -c * dt; s[3] = a * dt;
        return;
    }
    double r1 = 1. / s[0];
    double r3 = r1 * s[1];
    double r6 = 1. / (s[2] * r3 - s[3]);
    s[1] = r3 * r6;
    s[2] = r6 * s[2] * r1;
    s[0] = r1 - r3 * s[2];
    s[3] = -r6;
}

// utilities/matrixops.cpp:681-702
void invert3x3(double *m)
{
    double m00 = m[0], m01 = m[1], m02 = m[2], m10 = m[3], m11 = m[4], m12 = m[5], m20 = m[6], m21 = m[7], m22 = m[8];
    double r11 = std::fma(m11, m22, -m12 * m21);
    double r21 = std::fma(m12, m20, -m10 * m22);
    double r31 = std::fma(m10, m21, -m11 * m20);
    double det = 1. / std::fma(m00, r11, std::fma(m01, r21, m02 * r31));
    double v[9] = {r11, std::fma(m02, m21, -m01 * m22), std::fma(m01, m12, -m02 * m11),
                   r21, std::fma(m00, m22, -m02 * m20), std::fma(m02, m10, -m00 * m12),
                   r31, std::fma(m01, m20, -m00 * m21), std::fma(m11, m00, -m01 * m10)};
    for (int i = 0; i < 9; i++) m[i] = v[i] * det;
}

// utilities/matrixops.cpp:826-847
double det2(const double *s) { return std::fma(s[0], s[3], -s[1] * s[2]); }
double det3(const double *s)
{
    return std::fma(s[0], s[4] * s[8],
                    std::fma(s[5], s[6] * s[1],
                             std::fma(s[2], s[7] * s[3],
                                      std::fma(-s[0], s[5] * s[7],
                                               std::fma(-s[1], s[3] * s[8], -s[6] * s[4] * s[2])))));
}

// elements/integrable_entity.cpp:641-720: Jinv at reference point p.
void inverse_jacobian(int etype, const double *coords, double px, double py, double pz, double *Jinv)
{
    Shape sh = shape_of(etype);
    double dxi[6], deta[6], dzeta[6];
    dshape(etype, px, py, pz, dxi, deta, dzeta);
    if (sh.dim == 2) {
        double J[4] = {0., 0., 0., 0.};
        for (int i = 0; i < sh.nper; i++) {
            J[0] += dxi[i] * coords[2 * i];
            J[1] += dxi[i] * coords[2 * i + 1];
            J[2] += deta[i] * coords[2 * i];
            J[3] += deta[i] * coords[2 * i + 1];
        }
        invert2x2(J);
        std::memcpy(Jinv, J, sizeof(J));
    } else {
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < sh.nper; i++) {
            J[0] += dxi[i] * coords[3 * i];
            J[1] += dxi[i] * coords[3 * i + 1];
            J[2] += dxi[i] * coords[3 * i + 2];
            J[3] += deta[i] * coords[3 * i];
            J[4] += deta[i] * coords[3 * i + 1];
            J[5] += deta[i] * coords[3 * i + 2];
            J[6] += dzeta[i] * coords[3 * i];
            J[7] += dzeta[i] * coords[3 * i + 1];
            J[8] += dzeta[i] * coords[3 * i + 2];
        }
        invert3x3(J);
        std::memcpy(Jinv, J, sizeof(J));
    }
}

// Gauss rules: elements/elements.cpp:676-689 (P1 1-pt, P2 4-pt with the
// negative centroid weight) and :2399-2402 (Tet4 1-pt); weights divided by
// det(Jinv) at the point (:939-945, :2465-2470).
int gauss_points(int etype, double *pts, double *w)
{
    if (etype == 1) {
        pts[0] = 0.333333333333333; pts[1] = 0.333333333333333; pts[2] = 0.;
        w[0] = 0.5;
        return 1;
    }
    if (etype == 2) {
        const double P[4][2] = {{0.2, 0.2}, {0.6, 0.2}, {0.2, 0.6}, {0.33333333333333, 0.33333333333333}};
        const double W[4] = {0.260416666666667, 0.260416666666667, 0.260416666666667, -0.28125};
        for (int g = 0; g < 4; g++) { pts[3 * g] = P[g][0]; pts[3 * g + 1] = P[g][1]; pts[3 * g + 2] = 0.; w[g] = W[g]; }
        return 4;
    }
    pts[0] = 0.25; pts[1] = 0.25; pts[2] = 0.25;
    w[0] = 0.1666666666666667;
    return 1;
}

// polynomial/vm_base.cpp:2229-2256 (2D) and :2188-2228 (3D): B rows of
// node i (transpose = false, b x s) or its transpose (s x b).
void gradient(int dim, double dxi, double deta, double dzeta, const double *m, bool transpose, double *B)
{
    if (dim == 2) {
        double a = dxi * m[0] + deta * m[1];
        double c = dxi * m[2] + deta * m[3];
        if (!transpose) { // 2 x 3
            double t[6] = {a, 0., c, 0., c, a};
            std::memcpy(B, t, sizeof(t));
        } else { // 3 x 2
            double t[6] = {a, 0., 0., c, c, a};
            std::memcpy(B, t, sizeof(t));
        }
    } else {
        double a = dxi * m[0] + deta * m[1] + dzeta * m[2];
        double c = dxi * m[3] + deta * m[4] + dzeta * m[5];
        double e = dxi * m[6] + deta * m[7] + dzeta * m[8];
        if (!transpose) { // 3 x 6
            double t[18] = {a, 0., 0., 0., e, c,
                            0., c, 0., e, 0., a,
                            0., 0., e, c, a, 0.};
            std::memcpy(B, t, sizeof(t));
        } else { // 6 x 3
            double t[18] = {a, 0., 0.,
                            0., c, 0.,
                            0., 0., e,
                            0., e, c,
                            e, 0., a,
                            c, a, 0.};
            std::memcpy(B, t, sizeof(t));
        }
    }
}

// VirtualMachine::ieval(GtMtG) polynomial/vm_base.cpp:2974-2989 with
// matrix_matrix_matrix_multiply (utilities/matrixops.h:387-404) for the first
// Gauss point followed by `ret *= w`, and the Kahan
// matrix_matrix_matrix_multiply_and_add (utilities/matrixops.h:462-507) for
// the others.  Output: b x b block K_ij.
void ieval_block(int etype, const double *Jinv, int ngp, const double *gp, const double *w,
                 int i, int j, const double *C, double *ret)
{
    Shape sh = shape_of(etype);
    int b = sh.b, s = (b == 2) ? 3 : 6;
    double B[18], Bt[18], dxi[6], deta[6], dzeta[6];
    for (int q = 0; q < b * b; q++) ret[q] = 0.;
    for (int g = 0; g < ngp; g++) {
        dshape(etype, gp[3 * g], gp[3 * g + 1], gp[3 * g + 2], dxi, deta, dzeta);
        gradient(sh.dim, dxi[j], deta[j], dzeta[j], Jinv, true, Bt);   // third: Gradient(p_j, true)
        gradient(sh.dim, dxi[i], deta[i], dzeta[i], Jinv, false, B);   // first: Gradient(p_i)
        if (g == 0) {
            for (int r = 0; r < b; r++)
                for (int c = 0; c < s; c++) {
                    double rij = 0.;
                    for (int l = 0; l < s; l++) rij = rij + B[r * s + l] * C[l * s + c];
                    for (int k = 0; k < b; k++) ret[r * b + k] += rij * Bt[c * b + k];
                }
            for (int q = 0; q < b * b; q++) ret[q] *= w[g];
        } else {
            double comp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int r = 0; r < b; r++)
                for (int c = 0; c < s; c++) {
                    double rij = 0.;
                    for (int l = 0; l < s; l++) rij = rij + B[r * s + l] * C[l * s + c];
                    for (int k = 0; k < b; k++) {
                        double y = rij * Bt[c * b + k] * w[g] - comp[r * b + k];
                        double t = ret[r * b + k] + y;
                        comp[r * b + k] = (t - ret[r * b + k]) - y;
                        ret[r * b + k] = t;
                    }
                }
        }
    }
}

int col_len(int b) { return b + b % 2; }

// Pointer to block (row r, col c): sparse/sparse_vector.cpp:50-58 (binary
// search in the sorted column list of the block row).
double *block_ptr(int b, const uint32_t *row_start, const uint32_t *row_size, const uint32_t *col,
                  double *array, int64_t r, uint32_t c)
{
    const uint32_t *s = col + row_start[r], *e = s + row_size[r];
    const uint32_t *p = std::lower_bound(s, e, c);
    if (p == e || *p != c) return nullptr;
    return array + (size_t)(p - col) * b * col_len(b);
}

std::vector<uint32_t> row_starts(int64_t nrows, const uint32_t *row_size)
{
    std::vector<uint32_t> s(nrows + 1, 0);
    for (int64_t i = 0; i < nrows; i++) s[i + 1] = s[i] + row_size[i];
    return s;
}

// One block row of the Kahan SpMV, sparse/sparse_matrix.h:235-333: dest
// starts with its current content (-b or 0) and accumulates the row.
void spmv_row(int b, const double *array, const uint32_t *col, uint32_t start, uint32_t len,
              const double *v, double *dest)
{
    int cl = col_len(b);
    double comp[3] = {0., 0., 0.};
    for (uint32_t j = start; j < start + len; j++) {
        const double *a = array + (size_t)j * b * cl;
        const double *x = v + (size_t)col[j] * b;
        for (int c = 0; c < b; c++)
            for (int r = 0; r < b; r++) {
                double toAdd = a[c * cl + r] * x[c] - comp[r];
                double tot = dest[r] + toAdd;
                comp[r] = (tot - dest[r]) - toAdd;
                dest[r] = tot;
            }
    }
}

struct Csr {
    int b;
    int64_t nrows;
    const uint32_t *rs, *ci;
    const double *a;
    std::vector<uint32_t> start;
    Csr(int b_, int64_t n, const uint32_t *rs_, const uint32_t *ci_, const double *a_)
        : b(b_), nrows(n), rs(rs_), ci(ci_), a(a_), start(row_starts(n, rs_)) {}
    int64_t ndof() const { return nrows * b; }
    // y = A x - bvec (bvec nullable)
    void apply(const double *x, const double *bvec, double *y) const
    {
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < nrows; r++) {
            for (int k = 0; k < b; k++) y[r * b + k] = bvec ? -bvec[r * b + k] : 0.;
            spmv_row(b, a, ci, start[r], rs[r], x, y + r * b);
        }
    }
    double diag(int64_t d) const
    {
        int64_t r = d / b;
        int k = (int)(d % b);
        const uint32_t *s = ci + start[r], *e = s + rs[r];
        const uint32_t *p = std::lower_bound(s, e, (uint32_t)r);
        if (p == e || *p != (uint32_t)r) return 0.;
        return a[(size_t)(p - ci) * b * col_len(b) + k * col_len(b) + k];
    }
    double entry(int64_t i, int64_t j) const
    {
        int64_t r = i / b;
        const uint32_t *s = ci + start[r], *e = s + rs[r];
        const uint32_t *p = std::lower_bound(s, e, (uint32_t)(j / b));
        if (p == e || *p != (uint32_t)(j / b)) return 0.;
        return a[(size_t)(p - ci) * b * col_len(b) + (j % b) * col_len(b) + (i % b)];
    }
};

double kahan_dot(const double *v0, const double *v1, int64_t n)
{
    // utilities/matrixops.h:246-262
    double result = 0., compensate = 0.;
#pragma omp parallel for reduction(+ : result, compensate)
    for (int64_t i = 0; i < n; ++i) {
        double y = v0[i] * v1[i] - compensate;
        double tot = result + y;
        compensate = (tot - result) - y;
        result = tot;
    }
    return result;
}

} // namespace

extern "C" {

void or_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int or_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_cauchy_green(double p1, double p2, int dim, int plane_strain, double *C)
{
    // utilities/tensor.cpp:1845-1931, YOUNG_POISSON branch
    double Ciiii, Ciijj, Cijij;
    if (dim == 3 || plane_strain) {
        double c = p1 / ((1. + p2) * (1. - 2. * p2));
        Ciiii = c * (1. - p2); Ciijj = c * p2; Cijij = c * (1. - 2. * p2);
    } else {
        double c = p1 / (1. - p2 * p2);
        Ciiii = c; Ciijj = c * p2; Cijij = c * (1. - p2);
    }
    int s = dim == 3 ? 6 : 3;
    for (int i = 0; i < s * s; i++) C[i] = 0.;
    if (dim == 3) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) C[i * 6 + j] = (i == j) ? Ciiii : Ciijj;
        C[3 * 6 + 3] = C[4 * 6 + 4] = C[5 * 6 + 5] = Cijij;
    } else {
        C[0] = Ciiii; C[1] = Ciijj; C[3] = Ciijj; C[4] = Ciiii; C[8] = Cijij;
    }
}

void or_orient(int etype, const double *c, int *perm)
{
    Shape sh = shape_of(etype);
    for (int i = 0; i < sh.nper; i++) perm[i] = i;
    if (etype == 1 || etype == 2) {
        // geometry/geometry_2D.cpp:476-497 + geometry_base.cpp:3292-3307:
        // CCW test on consecutive edges of the vertex triangle; swap v1/v2.
        int v[3] = {0, etype == 2 ? 2 : 1, etype == 2 ? 4 : 2};
        bool ccw = true;
        for (int i = 0; i < 2; i++) {
            const double *p0 = c + 2 * v[i], *p1 = c + 2 * v[(i + 1) % 3], *p2 = c + 2 * v[(i + 2) % 3];
            double ax = p1[0] - p0[0], ay = p1[1] - p0[1];
            double bx = p2[0] - p1[0], by = p2[1] - p1[1];
            if (ax * by - ay * bx < 0) ccw = false;
        }
        if (!ccw) {
            if (etype == 1) { perm[1] = 2; perm[2] = 1; }
            else {
                // vertices v1 <-> v2; mid-nodes follow: m01 <-> m20, m12 fixed
                perm[2] = 4; perm[4] = 2; perm[1] = 5; perm[5] = 1;
            }
        }
    } else {
        // geometry/geometry_3D.cpp:254-268 with computeVolume :571-576
        const double *p0 = c, *p1 = c + 3, *p2 = c + 6, *p3 = c + 9;
        double u[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
        double v[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        double w[3] = {p1[0] - p3[0], p1[1] - p3[1], p1[2] - p3[2]};
        double x[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        double vol = (x[0] * w[0] + x[1] * w[1] + x[2] * w[2]) / 6.;
        if (vol < 0) { perm[0] = 1; perm[1] = 0; }
    }
}

void or_element_matrix(int etype, const double *coords, const double *C, double *Ke)
{
    Shape sh = shape_of(etype);
    double Jinv[9], gp[12], w[4];
    // Jinv once at the centroid (mesher/delaunay.cpp:2436-2447,
    // mesher/delaunay_3d.cpp:2478)
    if (sh.dim == 2) inverse_jacobian(etype, coords, 1. / 3., 1. / 3., 0., Jinv);
    else inverse_jacobian(etype, coords, .25, .25, .25, Jinv);
    int ngp = gauss_points(etype, gp, w);
    for (int g = 0; g < ngp; g++) {
        double Jg[9];
        inverse_jacobian(etype, coords, gp[3 * g], gp[3 * g + 1], gp[3 * g + 2], Jg);
        w[g] /= (sh.dim == 2) ? det2(Jg) : det3(Jg);
    }
    int b = sh.b, n = sh.nper * b;
    double blk[9];
    // mesher/delaunay.cpp:2457-2468: K_ii, then K_ij (j < i) and K_ji = K_ij^T
    for (int i = 0; i < sh.nper; i++) {
        for (int j = 0; j <= i; j++) {
            ieval_block(etype, Jinv, ngp, gp, w, i, j, C, blk);
            for (int r = 0; r < b; r++)
                for (int c = 0; c < b; c++) {
                    Ke[(i * b + r) * n + j * b + c] = blk[r * b + c];
                    if (j != i) Ke[(j * b + c) * n + i * b + r] = blk[r * b + c];
                }
        }
    }
}

int64_t or_pattern(int etype, int64_t nelem, const int32_t *conn, int64_t nnodes, uint32_t *row_size,
                   uint32_t *column_index)
{
    // solvers/assembly.cpp:448-524: the sorted set of (ids[j], ids[k]) node
    // pairs over each element, all diagonals (i,i) for i < nnodes, and the
    // (ndofmax-1, ndofmax-1) entry (ndofmax = nnodes, features.cpp:3381).
    Shape sh = shape_of(etype);
    std::vector<uint64_t> keys;
    keys.reserve((size_t)nelem * sh.nper * sh.nper + nnodes);
    for (int64_t e = 0; e < nelem; e++)
        for (int j = 0; j < sh.nper; j++)
            for (int k = 0; k < sh.nper; k++)
                keys.push_back(((uint64_t)(uint32_t)conn[e * sh.nper + j] << 32) | (uint32_t)conn[e * sh.nper + k]);
    for (int64_t i = 0; i < nnodes; i++) keys.push_back(((uint64_t)i << 32) | (uint64_t)i);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (row_size) {
        std::fill(row_size, row_size + nnodes, 0u);
        for (size_t q = 0; q < keys.size(); q++) {
            row_size[keys[q] >> 32]++;
            column_index[q] = (uint32_t)(keys[q] & 0xffffffffu);
        }
    }
    return (int64_t)keys.size();
}

void or_assemble(int etype, int64_t nelem, const int32_t *conn, const double *coords, const double *cmat,
                 const int32_t *phase, const double *damage, int64_t nrows, const uint32_t *row_size,
                 const uint32_t *column_index, double *array)
{
    or_assemble_dofmap(etype, nelem, conn, nullptr, coords, cmat, phase, damage, nrows, row_size, column_index, array);
}

void or_assemble_dofmap(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap, const double *coords,
                        const double *cmat, const int32_t *phase, const double *damage, int64_t nrows,
                        const uint32_t *row_size, const uint32_t *column_index, double *array)
{
    or_assemble_fractured(etype, nelem, conn, dofmap, coords, cmat, phase, damage, HUGE_VAL, 0., nrows, row_size,
                          column_index, array);
}

void or_assemble_fractured(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap, const double *coords,
                           const double *cmat, const int32_t *phase, const double *damage, double fractured_at,
                           double residual, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
                           double *array)
{
    Shape sh = shape_of(etype);
    int b = sh.b, cl = col_len(b), s = (b == 2) ? 3 : 6, n = sh.nper * b;
    std::vector<uint32_t> start = row_starts(nrows, row_size);
    size_t nnzb = start[nrows];
    std::fill(array, array + nnzb * b * cl, 0.);
    std::vector<double> comp(nnzb * b * cl, 0.);
    std::vector<double> Ke(n * n), xy(sh.nper * sh.dim), C(s * s);
    for (int64_t e = 0; e < nelem; e++) {
        const int32_t *geo = conn + e * sh.nper;
        int32_t ids[6];
        for (int i = 0; i < sh.nper; i++) ids[i] = dofmap ? dofmap[geo[i]] : geo[i];
        for (int i = 0; i < sh.nper; i++)
            for (int d = 0; d < sh.dim; d++) xy[i * sh.dim + d] = coords[(size_t)geo[i] * sh.dim + d];
        const double *Cp = cmat + (size_t)(phase ? phase[e] : 0) * s * s;
        // physics/damagemodels/isotropiclineardamage.cpp:39-46: C * (1 - d);
        // a fractured element keeps C * residual
        // (fiberbasedisotropiclineardamage.cpp:37-46)
        const double f = damage ? (damage[e] >= fractured_at ? residual : 1. - damage[e]) : 1.;
        for (int q = 0; q < s * s; q++) C[q] = damage ? Cp[q] * f : Cp[q];
        or_element_matrix(etype, xy.data(), C.data(), Ke.data());
        // solvers/assembly.cpp:657-735: Kahan adds, (j,j) then (j,k),(k,j) k>j.
        auto kadd = [&](double *a, double *c, double y0) {
            double y = y0 - *c;
            double t = *a + y;
            *c = (t - *a) - y;
            *a = t;
        };
        for (int j = 0; j < sh.nper; j++) {
            double *a = block_ptr(b, start.data(), row_size, column_index, array, ids[j], (uint32_t)ids[j]);
            double *c = comp.data() + (a - array);
            for (int m = 0; m < b; m++)
                for (int nn = 0; nn < b; nn++) kadd(a + m * cl + nn, c + m * cl + nn, Ke[(j * b + nn) * n + j * b + m]);
            for (int k = j + 1; k < sh.nper; k++) {
                double *a0 = block_ptr(b, start.data(), row_size, column_index, array, ids[j], (uint32_t)ids[k]);
                double *a1 = block_ptr(b, start.data(), row_size, column_index, array, ids[k], (uint32_t)ids[j]);
                double *c0 = comp.data() + (a0 - array), *c1 = comp.data() + (a1 - array);
                for (int m = 0; m < b; m++)
                    for (int nn = 0; nn < b; nn++) {
                        kadd(a0 + m * cl + nn, c0 + m * cl + nn, Ke[(j * b + nn) * n + k * b + m]);
                        kadd(a1 + m * cl + nn, c1 + m * cl + nn, Ke[(k * b + nn) * n + j * b + m]);
                    }
            }
        }
    }
}

void or_macro_rhs(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap, const double *coords,
                  const double *cmat, const int32_t *phase, const double *damage, const double *macro,
                  int64_t nrows, double *f)
{
    Shape sh = shape_of(etype);
    int b = sh.b, s = (b == 2) ? 3 : 6, n = sh.nper * b;
    std::fill(f, f + nrows * b, 0.);
    std::vector<double> Ke(n * n), xy(sh.nper * sh.dim), C(s * s), u(n);
    for (int64_t e = 0; e < nelem; e++) {
        const int32_t *geo = conn + e * sh.nper;
        for (int i = 0; i < sh.nper; i++)
            for (int d = 0; d < sh.dim; d++) xy[i * sh.dim + d] = coords[(size_t)geo[i] * sh.dim + d];
        const double *Cp = cmat + (size_t)(phase ? phase[e] : 0) * s * s;
        for (int q = 0; q < s * s; q++) C[q] = damage ? Cp[q] * (1. - damage[e]) : Cp[q];
        or_element_matrix(etype, xy.data(), C.data(), Ke.data());
        for (int i = 0; i < sh.nper; i++) {
            const double *X = &xy[i * sh.dim];
            if (sh.dim == 2) {
                u[i * 2] = macro[0] * X[0] + 0.5 * macro[2] * X[1];
                u[i * 2 + 1] = 0.5 * macro[2] * X[0] + macro[1] * X[1];
            } else {
                u[i * 3] = macro[0] * X[0] + 0.5 * macro[5] * X[1] + 0.5 * macro[4] * X[2];
                u[i * 3 + 1] = 0.5 * macro[5] * X[0] + macro[1] * X[1] + 0.5 * macro[3] * X[2];
                u[i * 3 + 2] = 0.5 * macro[4] * X[0] + 0.5 * macro[3] * X[1] + macro[2] * X[2];
            }
        }
        for (int i = 0; i < sh.nper; i++) {
            const int64_t row = dofmap ? dofmap[geo[i]] : geo[i];
            for (int r = 0; r < b; r++) {
                double t = 0.;
                for (int c = 0; c < n; c++) t += Ke[(i * b + r) * n + c] * u[c];
                f[row * b + r] -= t;
            }
        }
    }
}

void or_apply_dirichlet(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
                        double *array, double *rhs, int64_t nbc, const int64_t *dofs, const double *vals)
{
    // solvers/assembly.cpp:125-253: multipliers sorted by id; for every block
    // (k, col) of every block row, first the multipliers of row block k, then
    // those of column block col, in id order.
    std::vector<std::pair<int64_t, double>> mult(nbc);
    for (int64_t i = 0; i < nbc; i++) mult[i] = {dofs[i], vals[i]};
    std::stable_sort(mult.begin(), mult.end(), [](auto &x, auto &y) { return x.first < y.first; });
    // a later setting of the same dof replaces the earlier one
    // (Assembly::setPointAlong erases the duplicate, solvers/assembly.cpp:1590-1634)
    std::vector<std::pair<int64_t, double>> uniq;
    for (int64_t i = 0; i < nbc; i++) {
        if (!uniq.empty() && uniq.back().first == mult[i].first) uniq.back() = mult[i];
        else uniq.push_back(mult[i]);
    }
    mult.swap(uniq);
    nbc = (int64_t)mult.size();
    std::vector<int64_t> ids(nbc);
    for (int64_t i = 0; i < nbc; i++) ids[i] = mult[i].first;
    int cl = col_len(stride);
    std::vector<uint32_t> start = row_starts(nrows, row_size);
    for (int64_t k = 0; k < nrows; k++) {
        auto lo = [&](int64_t v) { return std::lower_bound(ids.begin(), ids.end(), v) - ids.begin(); };
        auto hi = [&](int64_t v) { return std::upper_bound(ids.begin(), ids.end(), v) - ids.begin(); };
        int64_t line_b = lo(k * stride), line_e = hi(k * stride + stride - 1);
        for (uint32_t l = 0; l < row_size[k]; l++) {
            int64_t cb = column_index[start[k] + l];
            double *blk = array + (size_t)(start[k] + l) * stride * cl;
            for (int pass = 0; pass < 2; pass++) {
                int64_t pb = pass == 0 ? line_b : lo(cb * stride);
                int64_t pe = pass == 0 ? line_e : hi(cb * stride + stride - 1);
                for (int64_t p = pb; p < pe; p++) {
                    int64_t id = mult[p].first;
                    double g = mult[p].second;
                    for (int m = 0; m < stride; m++) {
                        if (id != k * stride + m) {
                            for (int nn = 0; nn < stride; nn++)
                                if (id == cb * stride + nn) {
                                    double val = blk[cl * nn + m];
                                    rhs[k * stride + m] -= g * val;
                                    blk[cl * nn + m] = 0.;
                                }
                        } else {
                            if (pass == 1) rhs[id] = g;
                            for (int nn = 0; nn < stride; nn++) blk[cl * nn + m] = (cb * stride + nn == id) ? 1. : 0.;
                        }
                    }
                }
            }
        }
    }
}

void or_spmv(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index, const double *array,
             const double *x, const double *b, double *y)
{
    Csr A(stride, nrows, row_size, column_index, array);
    A.apply(x, b, y);
}

double or_dot(const double *a, const double *b, int64_t n) { return kahan_dot(a, b, n); }

void or_inverse_diagonal(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
                         const double *array, double *dinv)
{
    Csr A(stride, nrows, row_size, column_index, array);
    for (int64_t i = 0; i < A.ndof(); i++) {
        double v = A.diag(i);
        dinv[i] = (std::abs(v) > 1e-12) ? 1. / v : 0.;
    }
}

int or_inverse_2x2_diagonal(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
                            const double *array, double *blocks)
{
    Csr A(stride, nrows, row_size, column_index, array);
    int64_t n = A.ndof();
    if (n % 2) return -1; // the reference exits here (solvers/inversediagonal.cpp:86-90)
    for (int64_t i = 0; i < n; i += 2) {
        double s[4] = {A.entry(i, i), A.entry(i, i + 1), A.entry(i + 1, i), A.entry(i + 1, i + 1)};
        if (std::abs(det2(s)) > 1e-8) {
            invert2x2(s);
        } else {
            s[1] = 0.; s[2] = 0.;
            s[0] = std::abs(A.entry(i, i)) > 1e-8 ? 1. / A.entry(i, i) : 1.;
            s[3] = std::abs(A.entry(i + 1, i + 1)) > 1e-8 ? 1. / A.entry(i + 1, i + 1) : 1.;
        }
        std::memcpy(blocks + 2 * i, s, sizeof(s));
    }
    return 0;
}

int or_cg_reference(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
                    const double *array, const double *bvec, const double *x0, double eps, int64_t maxit,
                    int nssor, double *xout, int64_t *nit_out)
{
    // solvers/conjugategradient.cpp:69-318 (rowstart = colstart = 0)
    Csr A(stride, nrows, row_size, column_index, array);
    const int64_t n = A.ndof();
    *nit_out = 0;
    double bmax = 0.;
    for (int64_t i = 0; i < n; i++) bmax = std::max(bmax, std::abs(bvec[i]));
    std::vector<double> x(n, 0.), r(n), z(n), p(n), q(n), xmin(n), rc(n), xc(n), D(n);
    if (x0) std::copy(x0, x0 + n, x.begin());
    if (bmax < eps * eps) { // :74-78
        std::copy(x.begin(), x.end(), xout);
        return 1;
    }
    or_inverse_diagonal(stride, nrows, row_size, column_index, array, D.data());
    auto precond = [&](const std::vector<double> &v, std::vector<double> &t) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) t[i] = v[i] * D[i];
    };
    const double realeps = std::max(1e-12, eps);
    const int64_t Maxit = (maxit != -1) ? maxit : n / 2;
    int64_t nit = 0;
    double errmin = 1e9;
    xmin = x;
    int ret = 0;
    while (nit < Maxit) {
        int64_t localnit = 0;
        std::fill(rc.begin(), rc.end(), 0.);
        std::fill(xc.begin(), xc.end(), 0.);
        if (nssor) { // :128-152
            double err = 2, perr = 0;
            int64_t iter = 0;
            while (iter++ < nssor && err > realeps) {
                A.apply(x.data(), bvec, r.data());
                perr = err;
                err = std::sqrt(kahan_dot(r.data(), r.data(), n));
                if (err > perr) break;
                precond(r, r);
                for (int64_t i = 0; i < n; i++) {
                    double yx = -r[i] - xc[i];
                    double xt = x[i] + yx;
                    xc[i] = (xt - x[i]) - yx;
                    x[i] = xt;
                }
            }
            std::fill(xc.begin(), xc.end(), 0.);
        }
        A.apply(x.data(), bvec, r.data());
        for (int64_t i = 0; i < n; i++) r[i] *= -1;
        double err0 = std::sqrt(kahan_dot(r.data(), r.data(), n));
        if (std::isnan(err0)) { ret = -1; break; } // the reference exits the process
        if (nit == 0) errmin = err0;
        if (err0 < realeps) { ret = 1; break; }
        z = r;
        precond(r, z);
        p = z;
        A.apply(p.data(), nullptr, q.data());
        double last_rho = kahan_dot(r.data(), z.data(), n);
        double pq = kahan_dot(q.data(), p.data(), n);
        if (std::abs(pq) < 1e-12 * last_rho) { ret = 1; break; }
        double alpha = last_rho / pq;
        auto update = [&](double al) {
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; i++) {
                double yr = -q[i] * al - rc[i];
                double yx = p[i] * al - xc[i];
                double rt = r[i] + yr;
                double xt = x[i] + yx;
                rc[i] = (rt - r[i]) - yr;
                xc[i] = (xt - x[i]) - yx;
                r[i] = rt;
                x[i] = xt;
            }
        };
        update(alpha);
        double rho = 0, beta = 0;
        while (std::sqrt(std::abs(last_rho)) > realeps && localnit < n) { // :218-257
            localnit++;
            precond(r, z);
            rho = kahan_dot(r.data(), z.data(), n);
            beta = rho / last_rho;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; i++) p[i] = p[i] * beta + z[i];
            A.apply(p.data(), nullptr, q.data());
            pq = kahan_dot(q.data(), p.data(), n);
            if (std::abs(pq) < 1e-24 * rho) { last_rho = rho; break; }
            alpha = rho / pq;
            update(alpha);
            last_rho = rho;
            nit++;
        }
        A.apply(x.data(), bvec, r.data()); // :266-272
        double err = std::sqrt(kahan_dot(r.data(), r.data(), n));
        if (err < errmin) { errmin = std::sqrt(std::abs(rho)); xmin = x; }
        if (nssor) { // :274-302
            int64_t iters = 0;
            while (err > realeps && iters++ < n) {
                A.apply(x.data(), bvec, r.data());
                precond(r, r);
                for (int64_t i = 0; i < n; i++) {
                    double yx = -r[i] - xc[i];
                    double xt = x[i] + yx;
                    xc[i] = (xt - x[i]) - yx;
                    x[i] = xt;
                }
                double perr = err;
                err = std::sqrt(kahan_dot(r.data(), r.data(), n));
                if (err > perr) break;
                else xmin = x;
            }
            if (iters > 2) x = xmin;
        }
        if (std::min(err, std::sqrt(std::abs(last_rho))) < std::sqrt(realeps)) { ret = 1; break; } // :305
    }
    *nit_out = nit;
    std::copy(x.begin(), x.end(), xout);
    return ret;
}

int or_pcg(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index, const double *array,
           const double *bvec, const double *x0, double rtol, int64_t maxit, int precond_kind, double *xout,
           int64_t *nit_out, double *relres_out)
{
    Csr A(stride, nrows, row_size, column_index, array);
    const int64_t n = A.ndof();
    const int b = stride;
    std::vector<double> x(n, 0.), r(n), z(n), p(n), q(n), Minv;
    if (x0) std::copy(x0, x0 + n, x.begin());
    // preconditioner: point Jacobi or inverse diagonal blocks (row-major b x b)
    if (precond_kind == 0) {
        Minv.resize(n);
        for (int64_t i = 0; i < n; i++) { double d = A.diag(i); Minv[i] = std::abs(d) > 0 ? 1. / d : 0.; }
    } else {
        Minv.resize((size_t)nrows * b * b);
        for (int64_t k = 0; k < nrows; k++) {
            double m[9];
            for (int i = 0; i < b; i++)
                for (int j = 0; j < b; j++) m[i * b + j] = A.entry(k * b + i, k * b + j);
            if (b == 2) {
                double dt = 1. / (m[0] * m[3] - m[1] * m[2]);
                double t[4] = {m[3] * dt, -m[1] * dt, -m[2] * dt, m[0] * dt};
                std::memcpy(&Minv[(size_t)k * 4], t, sizeof(t));
            } else {
                double c[9] = {m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                               m[5] * m[6] - m[3] * m[8], m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                               m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
                double dt = 1. / (m[0] * c[0] + m[1] * c[3] + m[2] * c[6]);
                for (int i = 0; i < 9; i++) Minv[(size_t)k * 9 + i] = c[i] * dt;
            }
        }
    }
    auto prec = [&](const std::vector<double> &v, std::vector<double> &t) {
        if (precond_kind == 0) {
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; i++) t[i] = v[i] * Minv[i];
        } else {
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < nrows; k++)
                for (int i = 0; i < b; i++) {
                    double s = 0.;
                    for (int j = 0; j < b; j++) s += Minv[(size_t)k * b * b + i * b + j] * v[k * b + j];
                    t[k * b + i] = s;
                }
        }
    };
    auto dot = [&](const std::vector<double> &u, const std::vector<double> &v) {
        double s = 0.;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (int64_t i = 0; i < n; i++) s += u[i] * v[i];
        return s;
    };
    std::vector<double> bb(bvec, bvec + n);
    double bnorm = std::sqrt(dot(bb, bb));
    A.apply(x.data(), bvec, r.data());
    for (int64_t i = 0; i < n; i++) r[i] = -r[i];
    int64_t it = 0;
    double rnorm = std::sqrt(dot(r, r));
    if (bnorm == 0.) bnorm = 1.;
    int conv = rnorm <= rtol * bnorm;
    prec(r, z);
    p = z;
    double rho = dot(r, z);
    while (!conv && it < maxit) {
        A.apply(p.data(), nullptr, q.data());
        double alpha = rho / dot(p, q);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        it++;
        rnorm = std::sqrt(dot(r, r));
        if (rnorm <= rtol * bnorm) {
            // confirm with the true residual
            A.apply(x.data(), bvec, q.data());
            double tr = std::sqrt(dot(q, q));
            if (tr <= rtol * bnorm) { conv = 1; break; }
            for (int64_t i = 0; i < n; i++) r[i] = -q[i];
        }
        prec(r, z);
        double rho2 = dot(r, z);
        double beta = rho2 / rho;
        rho = rho2;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    A.apply(x.data(), bvec, q.data());
    *relres_out = std::sqrt(dot(q, q)) / bnorm;
    *nit_out = it;
    std::copy(x.begin(), x.end(), xout);
    return conv;
}

void or_element_strain(int etype, int64_t nelem, const int32_t *conn, const double *coords, const double *u,
                       double *strain)
{
    or_element_strain_dofmap(etype, nelem, conn, nullptr, coords, u, nullptr, strain);
}

// Strain and stress at every Gauss point of the element's own rule
// (gauss_points): ElementState::getField(TOTAL_STRAIN_FIELD, ...) at a
// local point (elements/integrable_entity.cpp:977-1097; the inverse
// Jacobian is the cache made at the centroid, :1011-1012) and
// REAL_STRESS_FIELD (:1379-1394: strain . C(1 - d)).  strain, stress:
// nelem x ngp x 3|6; stress may be null.
void or_gauss_fields(int etype, int64_t nelem, const int32_t *conn, const double *coords, const double *cmat,
                     const int32_t *phase, const double *damage, const double *u, double *strain, double *stress)
{
    Shape sh = shape_of(etype);
    const int s = sh.dim == 2 ? 3 : 6;
    double gp[12], gw[4];
    const int ngp = gauss_points(etype, gp, gw);
    for (int64_t e = 0; e < nelem; e++) {
        double xy[18], Jinv[9], dxi[6], deta[6], dzeta[6];
        const int32_t *ids = conn + e * sh.nper;
        for (int i = 0; i < sh.nper; i++)
            for (int d = 0; d < sh.dim; d++) xy[i * sh.dim + d] = coords[(size_t)ids[i] * sh.dim + d];
        const double cx = sh.dim == 2 ? 1. / 3. : .25, cz = sh.dim == 2 ? 0. : .25;
        inverse_jacobian(etype, xy, cx, cx, cz, Jinv);
        for (int g = 0; g < ngp; g++) {
            dshape(etype, gp[3 * g], gp[3 * g + 1], gp[3 * g + 2], dxi, deta, dzeta);
            double *out = strain + (e * ngp + g) * s;
            if (sh.dim == 2) {
                double x_xi = 0, x_eta = 0, y_xi = 0, y_eta = 0;
                for (int j = 0; j < sh.nper; j++) {
                    const double ux = u[(size_t)ids[j] * 2], uy = u[(size_t)ids[j] * 2 + 1];
                    x_xi += dxi[j] * ux; x_eta += deta[j] * ux;
                    y_xi += dxi[j] * uy; y_eta += deta[j] * uy;
                }
                out[0] = x_xi * Jinv[0] + x_eta * Jinv[1];
                out[1] = y_xi * Jinv[2] + y_eta * Jinv[3];
                out[2] = x_xi * Jinv[2] + x_eta * Jinv[3] + y_xi * Jinv[0] + y_eta * Jinv[1];
            } else {
                double gg[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int j = 0; j < sh.nper; j++)
                    for (int c = 0; c < 3; c++) {
                        const double v = u[(size_t)ids[j] * 3 + c];
                        gg[c][0] += dxi[j] * v; gg[c][1] += deta[j] * v; gg[c][2] += dzeta[j] * v;
                    }
                auto d = [&](int c, int row) { return gg[c][0] * Jinv[row * 3] + gg[c][1] * Jinv[row * 3 + 1] + gg[c][2] * Jinv[row * 3 + 2]; };
                out[0] = d(0, 0); out[1] = d(1, 1); out[2] = d(2, 2);
                out[3] = d(1, 2) + d(2, 1); out[4] = d(0, 2) + d(2, 0); out[5] = d(1, 0) + d(0, 1);
            }
            if (stress) {
                const double *Cp = cmat + (size_t)(phase ? phase[e] : 0) * s * s;
                const double f = damage ? 1. - damage[e] : 1.;
                for (int j = 0; j < s; j++) {
                    double t = 0.;
                    for (int i = 0; i < s; i++) t += out[i] * (Cp[i * s + j] * f);
                    stress[(e * ngp + g) * s + j] = t;
                }
            }
        }
    }
}

void or_element_strain_dofmap(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap,
                              const double *coords, const double *u, const double *macro, double *strain)
{
    Shape sh = shape_of(etype);
    int s = sh.dim == 2 ? 3 : 6;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nelem; e++) {
        double xy[18], Jinv[9], dxi[6], deta[6], dzeta[6];
        const int32_t *ids = conn + e * sh.nper;
        for (int i = 0; i < sh.nper; i++)
            for (int d = 0; d < sh.dim; d++) xy[i * sh.dim + d] = coords[(size_t)ids[i] * sh.dim + d];
        double px = sh.dim == 2 ? 1. / 3. : .25, py = px, pz = sh.dim == 2 ? 0. : .25;
        inverse_jacobian(etype, xy, px, py, pz, Jinv);
        dshape(etype, px, py, pz, dxi, deta, dzeta);
        double *out = strain + e * s;
        if (sh.dim == 2) {
            double x_xi = 0, x_eta = 0, y_xi = 0, y_eta = 0;
            for (int j = 0; j < sh.nper; j++) {
                const size_t dn = dofmap ? dofmap[ids[j]] : ids[j];
                double ux = u[dn * 2], uy = u[dn * 2 + 1];
                x_xi += dxi[j] * ux; x_eta += deta[j] * ux;
                y_xi += dxi[j] * uy; y_eta += deta[j] * uy;
            }
            out[0] = x_xi * Jinv[0] + x_eta * Jinv[1];
            out[1] = y_xi * Jinv[2] + y_eta * Jinv[3];
            out[2] = x_xi * Jinv[2] + x_eta * Jinv[3] + y_xi * Jinv[0] + y_eta * Jinv[1];
            if (macro) for (int q = 0; q < 3; q++) out[q] += macro[q];
        } else {
            double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}; // g[comp][ref dir]
            for (int j = 0; j < sh.nper; j++)
                for (int c = 0; c < 3; c++) {
                    double v = u[(size_t)(dofmap ? dofmap[ids[j]] : ids[j]) * 3 + c];
                    g[c][0] += dxi[j] * v; g[c][1] += deta[j] * v; g[c][2] += dzeta[j] * v;
                }
            auto d = [&](int c, int row) { return g[c][0] * Jinv[row * 3] + g[c][1] * Jinv[row * 3 + 1] + g[c][2] * Jinv[row * 3 + 2]; };
            out[0] = d(0, 0);
            out[1] = d(1, 1);
            out[2] = d(2, 2);
            out[3] = d(1, 2) + d(2, 1);
            out[4] = d(0, 2) + d(2, 0);
            out[5] = d(1, 0) + d(0, 1);
            if (macro) for (int q = 0; q < 6; q++) out[q] += macro[q];
        }
    }
}

int64_t or_damage_update(int64_t nelem, int dim, const double *strain, const double *d0, double kappa0,
                         double kappaf, double dmax, double *kappa, double *d)
{
    int64_t grown = 0;
#pragma omp parallel for reduction(+ : grown) schedule(static)
    for (int64_t e = 0; e < nelem; e++) {
        double ev[3];
        if (dim == 2) {
            const double *s = strain + e * 3;
            double a = s[0], bb = s[1], c = 0.5 * s[2];
            double m = 0.5 * (a + bb), r = std::sqrt(0.25 * (a - bb) * (a - bb) + c * c);
            ev[0] = m + r; ev[1] = m - r; ev[2] = 0.;
        } else {
            // symmetric tensor from Voigt [xx, yy, zz, yz, xz, xy] (engineering shear)
            const double *s = strain + e * 6;
            double a11 = s[0], a22 = s[1], a33 = s[2], a23 = 0.5 * s[3], a13 = 0.5 * s[4], a12 = 0.5 * s[5];
            double p1 = a12 * a12 + a13 * a13 + a23 * a23;
            double q = (a11 + a22 + a33) / 3.;
            if (p1 <= 1e-300) {
                ev[0] = a11; ev[1] = a22; ev[2] = a33;
            } else {
                double p2 = (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + (a33 - q) * (a33 - q) + 2. * p1;
                double p = std::sqrt(p2 / 6.);
                double b11 = (a11 - q) / p, b22 = (a22 - q) / p, b33 = (a33 - q) / p;
                double b12 = a12 / p, b13 = a13 / p, b23 = a23 / p;
                double detB = b11 * (b22 * b33 - b23 * b23) - b12 * (b12 * b33 - b23 * b13) + b13 * (b12 * b23 - b22 * b13);
                double rr = std::min(1., std::max(-1., detB / 2.));
                double phi = std::acos(rr) / 3.;
                ev[0] = q + 2. * p * std::cos(phi);
                ev[2] = q + 2. * p * std::cos(phi + 2. * M_PI / 3.);
                ev[1] = 3. * q - ev[0] - ev[2];
            }
        }
        double eq = 0.;
        for (int i = 0; i < 3; i++) { double t = std::max(0., ev[i]); eq += t * t; }
        eq = std::sqrt(eq);
        double k = std::max(kappa[e], eq);
        kappa[e] = k;
        double dl = 0.;
        if (k > kappa0) dl = 1. - (kappa0 / k) * std::exp(-(k - kappa0) / kappaf);
        double dn = std::min(dmax, std::max(d0 ? d0[e] : 0., std::max(d[e], dl)));
        if (dn > d[e]) grown++;
        d[e] = dn;
    }
    return grown;
}

void or_average_stress(int etype, int64_t nelem, const int32_t *conn, const double *coords, const double *cmat,
                       const int32_t *phase, const double *strain, double *out, double *measure_out)
{
    Shape sh = shape_of(etype);
    int s = sh.dim == 2 ? 3 : 6;
    std::vector<double> acc(s, 0.);
    double vol = 0.;
    for (int64_t e = 0; e < nelem; e++) {
        const int32_t *ids = conn + e * sh.nper;
        double meas;
        if (sh.dim == 2) {
            const double *a = coords + (size_t)ids[0] * 2, *b = coords + (size_t)ids[sh.nper == 6 ? 2 : 1] * 2,
                         *c = coords + (size_t)ids[sh.nper == 6 ? 4 : 2] * 2;
            meas = 0.5 * std::abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]));
        } else {
            const double *a = coords + (size_t)ids[0] * 3, *b = coords + (size_t)ids[1] * 3, *c = coords + (size_t)ids[2] * 3,
                         *d = coords + (size_t)ids[3] * 3;
            double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]},
                   w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
            meas = std::abs(u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0])) / 6.;
        }
        const double *C = cmat + (size_t)(phase ? phase[e] : 0) * s * s;
        for (int i = 0; i < s; i++) {
            double t = 0.;
            for (int j = 0; j < s; j++) t += C[i * s + j] * strain[e * s + j];
            acc[i] += meas * t;
        }
        vol += meas;
    }
    for (int i = 0; i < s; i++) out[i] = acc[i] / vol;
    if (measure_out) *measure_out = vol;
}

} // extern "C"

namespace {

// centroid, circumradius and measure of a linear element
struct ElemGeo {
    double c[3], r, meas;
};

ElemGeo elem_geo(int dim, int nper, const int32_t *ids, const double *coords)
{
    ElemGeo g{};
    double X[4][3] = {};
    for (int i = 0; i < nper; i++)
        for (int d = 0; d < dim; d++) X[i][d] = coords[(size_t)ids[i] * dim + d];
    for (int d = 0; d < 3; d++) {
        double s = 0.;
        for (int i = 0; i < nper; i++) s += X[i][d];
        g.c[d] = s / nper;
    }
    if (dim == 2) {
        // circumcentre: intersection of the perpendicular bisectors
        const double ax = X[0][0], ay = X[0][1], bx = X[1][0], by = X[1][1], cx = X[2][0], cy = X[2][1];
        const double dd = 2. * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by));
        const double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
        const double ux = (a2 * (by - cy) + b2 * (cy - ay) + c2 * (ay - by)) / dd;
        const double uy = (a2 * (cx - bx) + b2 * (ax - cx) + c2 * (bx - ax)) / dd;
        // Triangle::Triangle (geometry_2D.cpp:475-503): mean distance of the
        // first two points to the circumcentre
        g.r = (std::sqrt((bx - ux) * (bx - ux) + (by - uy) * (by - uy)) +
               std::sqrt((ax - ux) * (ax - ux) + (ay - uy) * (ay - uy))) / 2.;
        g.meas = 0.5 * std::abs((bx - ax) * (cy - ay) - (by - ay) * (cx - ax));
    } else {
        double u[3], v[3], w[3];
        for (int d = 0; d < 3; d++) { u[d] = X[1][d] - X[0][d]; v[d] = X[2][d] - X[0][d]; w[d] = X[3][d] - X[0][d]; }
        const double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
        // circumcentre relative to X0: solve 2 [u; v; w] x = [|u|^2, |v|^2, |w|^2]
        const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2],
                     ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        double vxw[3] = {v[1] * w[2] - v[2] * w[1], v[2] * w[0] - v[0] * w[2], v[0] * w[1] - v[1] * w[0]};
        double wxu[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
        double uxv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        double x[3];
        for (int d = 0; d < 3; d++) x[d] = (uu * vxw[d] + vv * wxu[d] + ww * uxv[d]) / (2. * det);
        // Tetrahedron::Tetrahedron (geometry_3D.cpp:254-281): distance of the
        // first point to the circumcentre
        g.r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        g.meas = std::abs(det) / 6.;
    }
    return g;
}

double smoothing_kernel(int kernel, double rrn)
{
    if (kernel == 1) return std::exp(-rrn);
    const double t = rrn / 2.;
    return 1. - t >= 0. ? (t - 1.) * (t - 1.) : 0.;
}

} // namespace

extern "C" {

void or_smoothed_strain(int etype, int64_t nelem, const int32_t *conn, const double *coords, const int32_t *phase,
                        const double *damage, double threshold, double radius, int kernel, const double *strain,
                        double *out)
{
    Shape sh = shape_of(etype);
    const int s = sh.dim == 2 ? 3 : 6;
    std::vector<ElemGeo> geo(nelem);
    for (int64_t e = 0; e < nelem; e++) geo[e] = elem_geo(sh.dim, sh.nper, conn + e * sh.nper, coords);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t e = 0; e < nelem; e++) {
        const ElemGeo &ge = geo[e];
        const double R = std::max(radius, ge.r * 1.1);
        const double rho = R * 2.5 + ge.r * 1.5;
        std::vector<double> acc(s, 0.), cor(s, 0.);
        double sum = 0.;
        for (int64_t i = 0; i < nelem; i++) {
            if (phase && phase[i] != phase[e]) continue;
            if (damage && damage[i] >= threshold) continue;
            const ElemGeo &gi = geo[i];
            double d2 = 0.;
            for (int d = 0; d < sh.dim; d++) d2 += (gi.c[d] - ge.c[d]) * (gi.c[d] - ge.c[d]);
            const bool in = sh.dim == 2 ? d2 < rho * rho : d2 < rho * rho + 2. * rho * 1e-12 + 1e-24;
            if (!in) continue;
            double coef = 0.;
            for (int v = 0; v < sh.nper; v++) {
                const double *X = coords + (size_t)conn[i * sh.nper + v] * sh.dim;
                double r2 = 0.;
                for (int d = 0; d < sh.dim; d++) r2 += (X[d] - ge.c[d]) * (X[d] - ge.c[d]);
                coef += std::max(smoothing_kernel(kernel, r2 / (R * R)) * gi.meas, 0.);
            }
            coef /= sh.nper;
            // getAverageField returns the element's measure times its weight
            // (elements/integrable_entity.cpp:2026-2133); the strains are
            // Kahan-added with it (mesher/mesh.h:1453-1462)
            const double v = gi.meas * coef;
            for (int q = 0; q < s; q++) {
                const double y = strain[i * s + q] * v - cor[q];
                const double t = acc[q] + y;
                cor[q] = (t - acc[q]) - y;
                acc[q] = t;
            }
            sum += v;
        }
        for (int q = 0; q < s; q++) out[e * s + q] = sum > 1e-12 ? acc[q] / sum : 0.;
    }
}


// toPrincipal (elements/integrable_entity.cpp:475-560): principal values of
// a stress (doubled = 0, SINGLE_OFF_DIAGONAL_VALUES) or an engineering
// strain (doubled = 1: the shear entries are 2 e_ij).  2D: (tr -/+ delta)/2
// in that order (cos of half the atan2 angle is never negative); 3D: the
// trigonometric cubic solution in the reference's operation order.
void or_principal(int dim, int doubled, const double *v, double *out)
{
    if (dim == 2) {
        const double tr = v[0] + v[1];
        const double det = doubled ? v[0] * v[1] - 0.25 * v[2] * v[2] : v[0] * v[1] - v[2] * v[2];
        const double delta = std::sqrt(tr * tr - 4. * det);
        const double angle = doubled ? 0.5 * std::atan2(0.5 * v[2], v[0] - v[1]) : 0.5 * std::atan2(v[2], v[0] - v[1]);
        if (std::cos(angle) < 0) {
            out[0] = (tr + delta) * .5;
            out[1] = (tr - delta) * .5;
        } else {
            out[0] = (tr - delta) * .5;
            out[1] = (tr + delta) * .5;
        }
        return;
    }
    // makeStressOrStrainMatrix (utilities/tensor.cpp): [xx xy xz; yx yy yz;
    // zx zy zz] from [xx yy zz yz xz xy], shear halved for strains
    const double h = doubled ? 0.5 : 1.;
    const double m00 = v[0], m11 = v[1], m22 = v[2], m12 = v[3] * h, m02 = v[4] * h, m01 = v[5] * h;
    const double trmat = -(m00 + m11 + m22);
    const double detmat = -2.0 * m01 * m02 * m12 + m00 * m12 * m12 + m11 * m02 * m02 + m22 * m01 * m01 - m00 * m11 * m22;
    const double m2mat = (m00 * m11 + m11 * m22 + m22 * m00) - m02 * m02 - m01 * m01 - m12 * m12;
    const double q = m2mat / 3. - trmat * trmat / 9.;
    const double r = (trmat * m2mat - 3. * detmat) / 6. - trmat * trmat * trmat / 27.;
    const double d = q * q * q + r * r;
    const double r0 = std::pow(r * r - d, 1. / 6.);
    double phi = std::atan2(std::sqrt(-1. * d), r) / 3.;
    if (std::abs(phi) < 1e-12) phi = 0.;
    if (phi < 0.) phi += M_PI;
    const double som = r0 * std::cos(phi), dif = r0 * std::sin(phi);
    out[0] = 2. * som - trmat / 3.;
    out[1] = -som - trmat / 3. - dif * std::sqrt(3.);
    out[2] = -som - trmat / 3. + dif * std::sqrt(3.);
}

// NonLocalMazars::grade (physics/fracturecriteria/mazars.cpp:24-31,
// 73-151) of the criterion each element holds -- a getCopy of
// NonLocalMazars(k0, E, nu, Gf, radius), made with the compressive
// parameters -1 (:200-205), hence tensionOnly false -- with
// FractureCriterion::step (fracturecriterion.cpp:534-560: a fractured
// element scores -1).  Per element of a phase with a criterion
// (crit[phase] != 0): the smoothed strain sbar (non-local, see
// or_smoothed_strain) gives the stress sbar.(C (1 - d)) and both principal
// sets;
//   kappa = gamma * max(0, sqrt(sum <e_i>+^2)), with e_3 = -nu (e1 + e2) /
//   (1 - nu) in plane stress; gamma != 1 only when every principal stress
//   is compressive;
//   plane stress: tension only (the 2D branch sets it, :94); 3D: tension
//   share t = sum <s_i - nu (s_j + s_k)>+ .. / (E (1 - d)) <e_i>+ / kappa^2
//   (0 below 1e-3), compression share 1 - t with A_c, B_c of cstress =
//   cstrain = -1;
//   pseudo = min(t (1 - (k0/K) exp(-B_t (K - k0))) + (1 - t) (1 - k0 (1 - A_c)/K
//            - A_c exp(-B_c (K - k0))), 1),  K = max(k0, kappa),
//   B_t = 0.1 E k0 / (Gf - 0.1 E k0^2 / 2),
//   score = pseudo - d.
// prm per phase: {k0, E, nu, Gf}.  Elements of a phase without a criterion
// score 0 and are never damaged.
void or_mazars_score(int dim, int64_t nelem, const int32_t *phase, const int32_t *crit, const double *prm,
                     const double *cmat, const double *sbar, const double *damage, double fractured_at, double *score)
{
    const int s = dim == 2 ? 3 : 6;
    for (int64_t e = 0; e < nelem; e++) {
        const int ph = phase ? phase[e] : 0;
        if (!crit[ph]) { score[e] = 0.; continue; }
        const double d = damage[e];
        if (d >= fractured_at) { score[e] = -1.; continue; }
        const double k0 = std::abs(prm[4 * ph]), E = prm[4 * ph + 1], nu = prm[4 * ph + 2], Gf = prm[4 * ph + 3];
        const double *eps = sbar + e * s;
        const double *Cp = cmat + (size_t)ph * s * s;
        double sig[6], ps[3] = {}, pe[3] = {};
        for (int j = 0; j < s; j++) {
            double t = 0.;
            for (int i = 0; i < s; i++) t += eps[i] * (Cp[i * s + j] * (1. - d));
            sig[j] = t;
        }
        or_principal(dim, 0, sig, ps);
        or_principal(dim, 1, eps, pe);
        const int np = dim == 2 ? 2 : 3;
        double trt = 0., trc = 0., sc2 = 0.;
        for (int i = 0; i < np; i++) {
            trt += 0.5 * (std::abs(ps[i]) + ps[i]);
            trc += std::min(0.0, ps[i]);
            sc2 += std::min(0.0, ps[i]) * std::min(0.0, ps[i]);
        }
        double gamma = 1.0;
        if (trc < -1e-12 && std::abs(trt / trc) < 1e-8) gamma = std::sqrt(sc2) / (-trc);
        double kappa, talpha;
        auto pos = [](double x) { return 0.5 * (std::abs(x) + x); };
        if (dim == 2) {
            const double ezz = -nu * (pe[0] + pe[1]) / (1. - nu);
            kappa = gamma * std::max(0.0, std::sqrt(std::pow(pos(pe[0]), 2.0) + std::pow(pos(pe[1]), 2.0) +
                                                    std::pow(pos(ezz), 2.0)));
            talpha = 1.0;
        } else {
            const double ed = E * (1 - d);
            const double q0 = (pos(ps[0]) - nu * (pos(ps[1]) + pos(ps[2]))) / ed;
            const double q1 = (pos(ps[1]) - nu * (pos(ps[0]) + pos(ps[2]))) / ed;
            const double q2 = (pos(ps[2]) - nu * (pos(ps[0]) + pos(ps[1]))) / ed;
            kappa = gamma * std::max(0., std::sqrt(std::pow(pos(pe[0]), 2.0) + std::pow(pos(pe[1]), 2.0) +
                                                   std::pow(pos(pe[2]), 2.0)));
            talpha = (q0 * pos(pe[0]) + q1 * pos(pe[1]) + q2 * pos(pe[2])) / (kappa * kappa);
            if (talpha < 1.0e-3) talpha = 0.;
        }
        const double calpha = 1.0 - talpha;
        const double cstress = -1., cstrain = -1.;
        const double Bt = (0.1 * E * k0) / (Gf - E * 0.1 * k0 * k0 * 0.5);
        const double Bc = -1. / (std::sqrt(2.) * nu * cstrain);
        const double Ac = -(E * k0 + std::sqrt(2.) * cstress * nu) / ((E * std::exp(Bc * k0 - 1.) / Bc) - E * k0);
        const double kt = std::max(k0, kappa);
        const double pseudo = std::min(talpha * (1. - (k0 / kt) * (std::exp(-Bt * (kt - k0)))) +
                                           calpha * (1. - (k0 * (1 - Ac) / kt) - Ac * (std::exp(-Bc * (kt - k0)))),
                                       1.0);
        score[e] = pseudo - d;
    }
}

// One pass of IsotropicLinearDamage::step, Newton branch (physics/
// damagemodels/isotropiclineardamage.cpp:64-92), with
// FractureCriterion::setChange at a checkpoint (fracturecriterion.cpp:
// 282-369; the Newton branch never clears the checkpoint) and postProcess
// (isotropiclineardamage.cpp:150-162).  T = the largest score of the pass
// over all criterion elements (FeatureTree's maxScoreInit).  Element e of a
// criterion phase is in its damaging set when T > 0, its score s > 0 and
// T - s <= 4 tol; its neighbourhood (the second cache, see or_fiber_pass)
// splits into that window and the rest, whose largest score m (max(0, T)
// when the rest is empty) gives delta = max(s - m, T - s).  Unless
// |T| < tol/2 and (e is out of the set or |s| < tol/20), an element in the
// set with s > tol/20 takes d += clamp(min(delta, s (1 - d)), 0, 1), capped
// at 1.  No decision reads damage written in the pass (postProcess adds it
// after every element has stepped), and the model never counts as
// fractured (its fraction is only set by DamageModel::step).  Returns the
// number of elements changed.
int64_t or_newton_pass(int etype, int64_t nelem, const int32_t *conn, const double *coords, const int32_t *phase,
                       const int32_t *crit, const double *score, double *damage, double radius, double tol,
                       int32_t *changed)
{
    Shape sh = shape_of(etype);
    std::vector<ElemGeo> geo(nelem);
    for (int64_t e = 0; e < nelem; e++) geo[e] = elem_geo(sh.dim, sh.nper, conn + e * sh.nper, coords);
    double T = -1.;
    for (int64_t e = 0; e < nelem; e++)
        if (crit[phase ? phase[e] : 0]) T = std::max(score[e], T);
    std::vector<double> inc(nelem, 0.);
    int64_t nch = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : nch)
    for (int64_t e = 0; e < nelem; e++) {
        changed[e] = 0;
        const int ph = phase ? phase[e] : 0;
        if (!crit[ph]) continue;
        const double s = score[e];
        const bool inset = T > 0 && T - s <= 4. * tol && s > 0.;
        double delta = 0.;
        if (inset) {
            const ElemGeo &ge = geo[e];
            const double R = std::max(radius, ge.r * 1.1);
            const double rho = std::max(R * 6., ge.r * 4.) * 2.5 + ge.r;
            double m = std::max(0., T);
            bool found = false;
            for (int64_t i = 0; i < nelem; i++) {
                if ((phase ? phase[i] : 0) != ph) continue;
                const ElemGeo &gi = geo[i];
                double d2 = 0.;
                for (int d = 0; d < sh.dim; d++) d2 += (gi.c[d] - ge.c[d]) * (gi.c[d] - ge.c[d]);
                const bool in = sh.dim == 2 ? d2 < rho * rho : d2 < rho * rho + 2. * rho * 1e-12 + 1e-24;
                if (!in) continue;
                if (T - score[i] <= 4. * tol && score[i] > 0.) continue;   // in the damaging window
                m = found ? std::max(score[i], m) : score[i];
                found = true;
            }
            delta = std::max(s - m, T - s);
        }
        if (std::abs(T) < .5 * tol && ((inset && std::abs(s) < .05 * tol) || !inset)) continue;
        if (s > .05 * tol && inset) {
            inc[e] = std::min(std::max(0. + std::min(delta, s * (1. - damage[e] - 0.)), 0.), 1.);
            changed[e] = 1;
            nch++;
        }
    }
    for (int64_t e = 0; e < nelem; e++)
        if (changed[e]) damage[e] = std::min(damage[e] + inc[e], 1.);
    return nch;
}

// One pass of FiberBasedIsotropicLinearDamage::step over the elements
// (physics/damagemodels/fiberbasedisotropiclineardamage.cpp:58-92, called
// for every element by FeatureTree::stepElements, features/features.cpp:
// 5047-5140): an element of a criterion phase that is not fractured, scores
// above 0 and is within tol (FractureCriterion::scoreTolerance, 2e-3) of the
// largest score in its neighbourhood takes d += fraction (capped at 1).  The
// neighbourhood is FractureCriterion::initialiseCache's second cache
// (fracturecriterion.cpp:171-190): same-phase elements whose centroid lies
// within max(6 R, 4 r) 2.5 + r of its centroid, R = max(radius, 1.1 r), r
// the circumradius; getMaxScoreInNeighbourhood (:251-280) takes the maximum
// over those that are not fractured when the element is visited, its own
// score included.  Elements are visited in the order the reference's pass
// runs them with one OpenMP thread: positions >= first_deferred, then
// 0 .. first_deferred-1 (libgomp runs the first 65 tasks of a single region
// last); damage written in the pass is seen by the elements visited after.
// first_deferred < 0: every element sees the damage from before the pass
// (an order-free pass, for comparison only).
// changed[e] = 1 for the elements incremented; nbr_max (nullable) receives
// the neighbourhood maximum each criterion element sees.  Returns the
// number incremented.
int64_t or_fiber_pass(int etype, int64_t nelem, const int32_t *conn, const double *coords, const int32_t *phase,
                      const int32_t *crit, const double *score, double *damage, double radius, double fraction,
                      double fractured_at, double tol, int64_t first_deferred, int32_t *changed, double *nbr_max)
{
    Shape sh = shape_of(etype);
    std::vector<ElemGeo> geo(nelem);
    for (int64_t e = 0; e < nelem; e++) geo[e] = elem_geo(sh.dim, sh.nper, conn + e * sh.nper, coords);
    int64_t nch = 0;
    const int64_t fd = std::min<int64_t>(std::max<int64_t>(first_deferred, 0), nelem);
    std::vector<double> before(damage, damage + nelem);
    const double *seen = first_deferred < 0 ? before.data() : damage;
    for (int64_t t = 0; t < nelem; t++) {
        const int64_t e = t < nelem - fd ? t + fd : t - (nelem - fd);
        changed[e] = 0;
        const int ph = phase ? phase[e] : 0;
        if (!crit[ph]) continue;
        if (!(score[e] > 0.) && !nbr_max) continue;
        const ElemGeo &ge = geo[e];
        const double R = std::max(radius, ge.r * 1.1);
        const double rho = std::max(R * 6., ge.r * 4.) * 2.5 + ge.r;
        double mx = score[e];
        for (int64_t i = 0; i < nelem; i++) {
            if ((phase ? phase[i] : 0) != ph) continue;
            const ElemGeo &gi = geo[i];
            double d2 = 0.;
            for (int d = 0; d < sh.dim; d++) d2 += (gi.c[d] - ge.c[d]) * (gi.c[d] - ge.c[d]);
            const bool in = sh.dim == 2 ? d2 < rho * rho : d2 < rho * rho + 2. * rho * 1e-12 + 1e-24;
            if (!in || seen[i] >= fractured_at) continue;
            mx = std::max(mx, score[i]);
        }
        if (nbr_max) nbr_max[e] = mx;
        if (score[e] > 0. && damage[e] < fractured_at && std::abs(score[e] - mx) < tol) {
            damage[e] = std::min(damage[e] + fraction, 1.);
            changed[e] = 1;
            nch++;
        }
    }
    return nch;
}

} // extern "C"

// ---------------------------------------------------------------- mesh input
int64_t or_renumber(int64_t nnodes, int64_t nelem, int nper, const int32_t *conn, int32_t *ids)
{
    // features/features.cpp:603-633: ids reset to -1, then every bounding
    // point of every element, in iteration order, gets count++ if still -1
    std::fill(ids, ids + nnodes, -1);
    int64_t count = 0;
    for (int64_t q = 0; q < nelem * nper; q++)
        if (ids[conn[q]] == -1) ids[conn[q]] = (int32_t)count++;
    return count;
}

int64_t or_add_midpoints(int64_t nnodes, const double *xy, int64_t ntri, const int32_t *conn3, int32_t *conn6,
                         double *mid_xy)
{
    // mesher/delaunay.cpp:280-351 with nodes_per_side = 1: fraction 0 finds
    // the side's first vertex, fraction 1/2 the mid-edge point, looked up in
    // the already visited neighbours (an edge's other triangle) by position;
    // a miss makes a new point a*(1 - 0.5) + b*0.5
    std::unordered_map<uint64_t, int32_t> edge;
    edge.reserve((size_t)ntri * 2);
    int64_t nmid = 0;
    for (int64_t t = 0; t < ntri; t++)
        for (int s = 0; s < 3; s++) {
            const int32_t a = conn3[t * 3 + s], b = conn3[t * 3 + (s + 1) % 3];
            const uint64_t key = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            auto it = edge.find(key);
            int32_t id;
            if (it != edge.end()) {
                id = it->second;
            } else {
                id = (int32_t)(nnodes + nmid);
                for (int d = 0; d < 2; d++) mid_xy[nmid * 2 + d] = xy[(int64_t)a * 2 + d] * (1. - .5) + xy[(int64_t)b * 2 + d] * .5;
                edge.emplace(key, id);
                nmid++;
            }
            conn6[t * 6 + 2 * s] = a;
            conn6[t * 6 + 2 * s + 1] = id;
        }
    return nmid;
}

// ---------------------------------------------------------------- voxels
namespace {
double det3(const double *a, const double *b, const double *c)
{
    return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

// barycentric test of geometry_3D.cpp:451-533 (weights of vertices 2, 3, 0),
// by volumes instead of the reference's 4x4 inverse
bool in_tet(const double (*P)[3], const double *p)
{
    double e[3][3], f[3][3];
    for (int d = 0; d < 3; d++) {
        e[0][d] = P[3][d] - P[2][d];
        e[1][d] = P[0][d] - P[2][d];
        e[2][d] = P[1][d] - P[2][d];
    }
    const double V = det3(e[0], e[1], e[2]);
    if (V == 0.) return false;
    double q[3][3];
    for (int d = 0; d < 3; d++) { q[0][d] = P[3][d] - p[d]; q[1][d] = P[0][d] - p[d]; q[2][d] = P[1][d] - p[d]; }
    const double l2 = det3(q[0], q[1], q[2]) / V;
    for (int d = 0; d < 3; d++) f[0][d] = p[d] - P[2][d];
    const double l3 = det3(f[0], e[1], e[2]) / V;
    const double l0 = det3(e[0], f[0], e[2]) / V;
    if (l2 < -1e-12 || l3 < -1e-12 || l0 < -1e-12) return false;
    return !(l2 + l3 + l0 > 1 + 3. * 1e-12);
}
} // namespace

void or_voxel_owner(int64_t nnodes, const double *xyz, int64_t ntet, const int32_t *conn, const double *box,
                    int nx, int ny, int nz, int32_t *owner)
{
    (void)nnodes;
    const int n[3] = {nx, ny, nz};
    double bl[3], tr[3];
    for (int d = 0; d < 3; d++) {
        const double v = box[3 + d] * (n[d] - 1) / (n[d]) * 0.5;
        bl[d] = box[d] - v;
        tr[d] = box[d] + v;
    }
    std::fill(owner, owner + (int64_t)nx * ny * nz, -1);
    for (int64_t t = 0; t < ntet; t++) {
        double P[4][3], lo[3], hi[3];
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) P[a][d] = xyz[(int64_t)conn[t * 4 + a] * 3 + d];
        for (int d = 0; d < 3; d++) {
            lo[d] = hi[d] = P[0][d];
            for (int a = 1; a < 4; a++) { lo[d] = std::min(lo[d], P[a][d]); hi[d] = std::max(hi[d], P[a][d]); }
        }
        for (int i = (int)(nx * (lo[0] - bl[0]) / (tr[0] - bl[0])); i < nx * (hi[0] - bl[0]) / (tr[0] - bl[0]); i++) {
            if (i < 0 || i >= nx) continue;
            for (int j = (int)(ny * (lo[1] - bl[1]) / (tr[1] - bl[1])); j < ny * (hi[1] - bl[1]) / (tr[1] - bl[1]); j++) {
                if (j < 0 || j >= ny) continue;
                for (int k = (int)(nz * (lo[2] - bl[2]) / (tr[2] - bl[2])); k < nz * (hi[2] - bl[2]) / (tr[2] - bl[2]); k++) {
                    if (k < 0 || k >= nz) continue;
                    double p[3] = {bl[0], bl[1], bl[2]};
                    p[0] += (tr[0] - bl[0]) * ((double)i) / (double)(nx - 1);
                    p[1] += (tr[1] - bl[1]) * ((double)j) / (double)(ny - 1);
                    p[2] += (tr[2] - bl[2]) * ((double)k) / (double)(nz - 1);
                    bool in = in_tet(P, p);
                    for (int m = 0; m < 8 && !in; m++) {
                        const double q[3] = {p[0] + ((m & 4) ? 16. : -16.) * 1e-12, p[1] + ((m & 2) ? 16. : -16.) * 1e-12,
                                             p[2] + ((m & 1) ? 16. : -16.) * 1e-12};
                        in = in_tet(P, q);
                    }
                    if (in) owner[k + (int64_t)ny * j + (int64_t)ny * nx * i] = (int32_t)t;
                }
            }
        }
    }
}

