// Element integration, gather and coloured assembly, Dirichlet elimination, scaling,
// strain / damage kernels (device side).
//
// Element matrices follow the reference quadrature exactly in structure:
//   Jinv once at the element centroid (mesher/delaunay.cpp:2436-2447,
//   mesher/delaunay_3d.cpp:2478), Gauss rules of elements/elements.cpp:676-689
//   and :2399-2402 with weights divided by det(Jinv) (:939-945), B rows of
//   polynomial/vm_base.cpp:2188-2256, K_ij = sum_g w_g B_i C B_j^T with the
//   first Gauss point as a plain product scaled by w and later points added
//   with Kahan compensation (polynomial/vm_base.cpp:2974-2989,
//   utilities/matrixops.h:462-507), K_ji = K_ij^T (mesher/delaunay.cpp:2457-2468).
// Two scatters into the FP64 matrix, both without atomics: the gather (one
// thread per block, its contributions in the reference's order) and the
// coloured one (one thread per element adding its blocks; elements of one
// colour share no DOF node, so the order per block is the colour order).
#pragma once
#include "device.hpp"

namespace amie {

#define AMIE_HD __host__ __device__
// The element formulas are evaluated without fused multiply-add contraction,
// in the reference's operation order: its 2x2 inverse
// (utilities/matrixops.cpp:536-561) branches on J[0][0] being exactly zero,
// which a contracted P2 Jacobian sum would turn into a ~1e-17 residue.
#define AMIE_NO_CONTRACT _Pragma("clang fp contract(off)")

struct AsmArgs {
    const double *coords;     // geometry coordinates, dim per node
    const int32_t *econn;     // element -> geometry node (global)
    const int32_t *erow;      // element node -> internal block row (global)
    const uint32_t *eslot;    // element (i,j) -> SELL slot (global)
    const int32_t *ephase;
    const double *edamage;    // nullable
    const int32_t *eprob;
    const int32_t *cmat_off;  // per problem: first C matrix
    const double *cmat;
    const double *macro;      // per problem nrhs x s (nullable)
    double *v64;
    int64_t nslots;
    double *b64;              // right-hand-side planes (vplane entries each)
    int64_t vplane;
    double frac_at = HUGE_VAL;   // damage at which an element counts as fractured (HUGE_VAL: never)
    double residual = 0.;        // stiffness fraction a fractured element keeps
};

// Stiffness factor of element e: C (1 - d) (physics/damagemodels/
// isotropiclineardamage.cpp:39-46), or C * residual once d >= frac_at
// (fiberbasedisotropiclineardamage.cpp:37-46).
AMIE_HD inline double stiffness_factor(const double *ed, int64_t e, double frac_at, double residual)
{
    return ed ? (ed[e] >= frac_at ? residual : 1. - ed[e]) : 1.;
}

template <int ET> struct Elem;
template <> struct Elem<1> { static constexpr int NPER = 3, DIM = 2, B = 2, S = 3, NGP = 1; };
template <> struct Elem<2> { static constexpr int NPER = 6, DIM = 2, B = 2, S = 3, NGP = 4; };
template <> struct Elem<3> { static constexpr int NPER = 4, DIM = 3, B = 3, S = 6, NGP = 1; };

// shape-function derivatives (see oracle: elements/elements.cpp:976-1061,2491-2530)
template <int ET> AMIE_HD inline void dshape(double x, double y, double *dxi, double *deta, double *dzeta)
{
    AMIE_NO_CONTRACT
    if constexpr (ET == 1) {
        dxi[0] = 0.; deta[0] = 1.;
        dxi[1] = -1.; deta[1] = -1.;
        dxi[2] = 1.; deta[2] = 0.;
    } else if constexpr (ET == 2) {
        dxi[0] = 0.;                     deta[0] = 4. * y - 1.;
        dxi[1] = y * -4.;                deta[1] = (4. - x * 4.) - y * 8.;
        dxi[2] = (y * 4. + x * 4.) - 3.; deta[2] = (x * 4. + y * 4.) - 3.;
        dxi[3] = (4. - 8. * x) - y * 4.; deta[3] = x * -4.;
        dxi[4] = 4. * x - 1.;            deta[4] = 0.;
        dxi[5] = 4. * y;                 deta[5] = 4. * x;
    } else {
        dxi[0] = 0.; deta[0] = 0.; dzeta[0] = 1.;
        dxi[1] = -1.; deta[1] = -1.; dzeta[1] = -1.;
        dxi[2] = 1.; deta[2] = 0.; dzeta[2] = 0.;
        dxi[3] = 0.; deta[3] = 1.; dzeta[3] = 0.;
    }
}

// Jinv at the centroid (utilities/matrixops.cpp:536-561, :681-702) and its
// determinant (:826-847).
template <int ET> AMIE_HD inline double centroid_jinv(const double (*X)[3], double *m)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    double dxi[6], deta[6], dzeta[6];
    if constexpr (E::DIM == 2) {
        dshape<ET>(1. / 3., 1. / 3., dxi, deta, dzeta);
        double J[4] = {0., 0., 0., 0.};
        for (int i = 0; i < E::NPER; i++) {
            J[0] += dxi[i] * X[i][0]; J[1] += dxi[i] * X[i][1];
            J[2] += deta[i] * X[i][0]; J[3] += deta[i] * X[i][1];
        }
        if (fabs(J[0]) < 1e-24) {
            double dt = 1. / (J[0] * J[3] - J[1] * J[2]);
            m[0] = J[3] * dt; m[1] = -J[1] * dt; m[2] = -J[2] * dt; m[3] = J[0] * dt;
        } else {
            double r1 = 1. / J[0], r3 = r1 * J[1], r6 = 1. / (J[2] * r3 - J[3]);
            m[1] = r3 * r6;
            m[2] = r6 * J[2] * r1;
            m[0] = r1 - r3 * m[2];
            m[3] = -r6;
        }
        return fma(m[0], m[3], -m[1] * m[2]);
    } else {
        dshape<ET>(.25, .25, dxi, deta, dzeta);
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < E::NPER; i++)
            for (int d = 0; d < 3; d++) {
                J[d] += dxi[i] * X[i][d]; J[3 + d] += deta[i] * X[i][d]; J[6 + d] += dzeta[i] * X[i][d];
            }
        double r11 = fma(J[4], J[8], -J[5] * J[7]);
        double r21 = fma(J[5], J[6], -J[3] * J[8]);
        double r31 = fma(J[3], J[7], -J[4] * J[6]);
        double det = 1. / fma(J[0], r11, fma(J[1], r21, J[2] * r31));
        double v[9] = {r11, fma(J[2], J[7], -J[1] * J[8]), fma(J[1], J[5], -J[2] * J[4]),
                       r21, fma(J[0], J[8], -J[2] * J[6]), fma(J[2], J[3], -J[0] * J[5]),
                       r31, fma(J[1], J[6], -J[0] * J[7]), fma(J[4], J[0], -J[1] * J[3])};
        for (int i = 0; i < 9; i++) m[i] = v[i] * det;
        return fma(m[0], m[4] * m[8],
                   fma(m[5], m[6] * m[1], fma(m[2], m[7] * m[3], fma(-m[0], m[5] * m[7], fma(-m[1], m[3] * m[8], -m[6] * m[4] * m[2])))));
    }
}

template <int ET> AMIE_HD inline void gauss(int g, double &x, double &y, double &w)
{
    if constexpr (ET == 1) { x = 0.333333333333333; y = x; w = 0.5; }
    else if constexpr (ET == 2) {
        const double P[4][2] = {{0.2, 0.2}, {0.6, 0.2}, {0.2, 0.6}, {0.33333333333333, 0.33333333333333}};
        x = P[g][0]; y = P[g][1];
        w = g < 3 ? 0.260416666666667 : -0.28125;
    } else { x = .25; y = .25; w = 0.1666666666666667; }
}

// Physical gradients G[g][node][dim] of the shape functions at every Gauss
// point (with the centroid Jinv) and the Gauss weights w[g].
template <int ET> AMIE_HD inline void element_grads(const double (*X)[3], double (*G)[6][3], double *w)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    double m[9];
    double detJinv = centroid_jinv<ET>(X, m);
    for (int g = 0; g < E::NGP; g++) {
        double gx, gy, gw, dxi[6], deta[6], dzeta[6];
        gauss<ET>(g, gx, gy, gw);
        dshape<ET>(gx, gy, dxi, deta, dzeta);
        w[g] = gw / detJinv;
        for (int i = 0; i < E::NPER; i++) {
            if constexpr (E::DIM == 2) {
                G[g][i][0] = dxi[i] * m[0] + deta[i] * m[1];
                G[g][i][1] = dxi[i] * m[2] + deta[i] * m[3];
            } else {
                G[g][i][0] = dxi[i] * m[0] + deta[i] * m[1] + dzeta[i] * m[2];
                G[g][i][1] = dxi[i] * m[3] + deta[i] * m[4] + dzeta[i] * m[5];
                G[g][i][2] = dxi[i] * m[6] + deta[i] * m[7] + dzeta[i] * m[8];
            }
        }
    }
}

// B_i C B_j^T for one Gauss point; out row-major b x b.
template <int ET> AMIE_HD inline void btcb(const double *gi, const double *gj, const double *C, double *out)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int S = E::S, B = E::B;
    double Bi[B][S], Bj[S][B];
    for (int r = 0; r < B; r++) for (int c = 0; c < S; c++) Bi[r][c] = 0.;
    for (int r = 0; r < S; r++) for (int c = 0; c < B; c++) Bj[r][c] = 0.;
    if constexpr (B == 2) {
        Bi[0][0] = gi[0]; Bi[0][2] = gi[1]; Bi[1][1] = gi[1]; Bi[1][2] = gi[0];
        Bj[0][0] = gj[0]; Bj[1][1] = gj[1]; Bj[2][0] = gj[1]; Bj[2][1] = gj[0];
    } else {
        Bi[0][0] = gi[0]; Bi[1][1] = gi[1]; Bi[2][2] = gi[2];
        Bi[1][3] = gi[2]; Bi[2][3] = gi[1]; Bi[0][4] = gi[2]; Bi[2][4] = gi[0]; Bi[0][5] = gi[1]; Bi[1][5] = gi[0];
        Bj[0][0] = gj[0]; Bj[1][1] = gj[1]; Bj[2][2] = gj[2];
        Bj[3][1] = gj[2]; Bj[3][2] = gj[1]; Bj[4][0] = gj[2]; Bj[4][2] = gj[0]; Bj[5][0] = gj[1]; Bj[5][1] = gj[0];
    }
    for (int r = 0; r < B; r++)
        for (int k = 0; k < B; k++) out[r * B + k] = 0.;
    for (int r = 0; r < B; r++)
        for (int c = 0; c < S; c++) {
            double rij = 0.;
            for (int l = 0; l < S; l++) rij = rij + Bi[r][l] * C[l * S + c];
            for (int k = 0; k < B; k++) out[r * B + k] += rij * Bj[c][k];
        }
}

// Dense element matrix (nper*b)^2, dof = local_node*b + axis: the exact
// formulation the assembly kernel adds into the global matrix.
// gi[g], gj[g]: gradients of local nodes i and j at Gauss point g.
template <int ET> AMIE_HD inline void element_block_pair(const double (*gi)[3], const double (*gj)[3], const double *w,
                                                         const double *C, double *K)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int B = E::B;
    double t[B * B], comp[B * B];
    btcb<ET>(gi[0], gj[0], C, K);
    for (int q = 0; q < B * B; q++) { K[q] *= w[0]; comp[q] = 0.; }
    for (int g = 1; g < E::NGP; g++) {
        btcb<ET>(gi[g], gj[g], C, t);
        for (int q = 0; q < B * B; q++) {
            double yk = t[q] * w[g] - comp[q];
            double tt = K[q] + yk;
            comp[q] = (tt - K[q]) - yk;
            K[q] = tt;
        }
    }
}

template <int ET> AMIE_HD inline void element_block(const double (*G)[6][3], const double *w, const double *C, int i,
                                                    int j, double *K)
{
    using E = Elem<ET>;
    double gi[E::NGP][3], gj[E::NGP][3];
    for (int g = 0; g < E::NGP; g++)
        for (int d = 0; d < 3; d++) { gi[g][d] = G[g][i][d]; gj[g][d] = G[g][j][d]; }
    element_block_pair<ET>(gi, gj, w, C, K);
}

// Gradients of local node n (a run-time index) picked from G with
// compile-time indexing only, so that G stays in registers.
template <int ET> __device__ inline void pick_grads(const double (*G)[6][3], int n, double (*gn)[3])
{
    using E = Elem<ET>;
    for (int m = 0; m < E::NPER; m++)
        for (int g = 0; g < E::NGP; g++)
            for (int d = 0; d < 3; d++)
                if (m == n) gn[g][d] = G[g][m][d];
}

template <int ET> AMIE_HD inline void element_matrix(const double (*X)[3], const double *C, double *Ke)
{
    using E = Elem<ET>;
    constexpr int B = E::B, NP = E::NPER, N = NP * B;
    double G[E::NGP][6][3], w[E::NGP];
    element_grads<ET>(X, G, w);
    for (int i = 0; i < NP; i++)
        for (int j = 0; j <= i; j++) {
            double K[B * B];
            element_block<ET>(G, w, C, i, j, K);
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++) {
                    Ke[(i * B + r) * N + j * B + c] = K[r * B + c];
                    Ke[(j * B + c) * N + i * B + r] = K[r * B + c];
                }
        }
}

// Element matrices of a list of elements (no assembly): nelem x (nper*b)^2.
template <int ET>
__global__ __launch_bounds__(256) void k_element_matrices(int n, const double *coords, const double *cmat, double *Ke)
{
    using E = Elem<ET>;
    constexpr int NP = E::NPER, D = E::DIM, S = E::S, N = NP * E::B;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double X[NP][3];
    for (int i = 0; i < NP; i++) {
        for (int d = 0; d < D; d++) X[i][d] = coords[((int64_t)e * NP + i) * D + d];
        if (D == 2) X[i][2] = 0.;
    }
    double C[S * S];
    for (int q = 0; q < S * S; q++) C[q] = cmat[(int64_t)e * S * S + q];
    element_matrix<ET>(X, C, Ke + (int64_t)e * N * N);
}

template <int B> __device__ inline void add_block64(double *v64, int64_t ns, int64_t slot, const double *blk /*row-major*/)
{
    // planes of double2 (entry (r,c) at c*B + r), remainder plane of double
    double a[B * B];
    for (int r = 0; r < B; r++)
        for (int c = 0; c < B; c++) a[c * B + r] = blk[r * B + c];
    double2 *p = reinterpret_cast<double2 *>(v64);
    constexpr int NP = (B * B) / 2;
    for (int k = 0; k < NP; k++) {
        double2 v = p[(int64_t)k * ns + slot];
        v.x += a[2 * k]; v.y += a[2 * k + 1];
        p[(int64_t)k * ns + slot] = v;
    }
    if constexpr ((B * B) % 2) {
        double *rem = reinterpret_cast<double *>(p + (int64_t)NP * ns);
        rem[slot] += a[B * B - 1];
    }
}

template <int ET, int NR>
__global__ __launch_bounds__(256) void k_assemble(AsmArgs a, const int32_t *elems, int n, int S_voigt)
{
    using E = Elem<ET>;
    constexpr int NP = E::NPER, B = E::B, S = E::S, D = E::DIM;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int e = elems[gid];
    const int p = a.eprob[e];
    double X[NP][3];
    for (int i = 0; i < NP; i++) {
        const int nd = a.econn[(int64_t)e * NP + i];
        for (int d = 0; d < D; d++) X[i][d] = a.coords[(int64_t)nd * D + d];
        if (D == 2) X[i][2] = 0.;
    }
    double C[S * S];
    const double *Cp = a.cmat + (int64_t)(a.cmat_off[p] + a.ephase[e]) * S * S;
    const double f = stiffness_factor(a.edamage, e, a.frac_at, a.residual);
    for (int k = 0; k < S * S; k++) C[k] = a.edamage ? Cp[k] * f : Cp[k];
    double G[E::NGP][6][3], w[E::NGP];
    element_grads<ET>(X, G, w);
    // macro-strain displacement of every node, per right-hand side
    double U[NR][NP][B];
    const bool macro = a.macro != nullptr;
    if (macro) {
        for (int k = 0; k < NR; k++) {
            const double *m = a.macro + ((int64_t)p * NR + k) * S;
            for (int i = 0; i < NP; i++) {
                if constexpr (D == 2) {
                    U[k][i][0] = m[0] * X[i][0] + 0.5 * m[2] * X[i][1];
                    U[k][i][1] = 0.5 * m[2] * X[i][0] + m[1] * X[i][1];
                } else {
                    // Voigt [xx, yy, zz, yz, xz, xy]
                    U[k][i][0] = m[0] * X[i][0] + 0.5 * m[5] * X[i][1] + 0.5 * m[4] * X[i][2];
                    U[k][i][1] = 0.5 * m[5] * X[i][0] + m[1] * X[i][1] + 0.5 * m[3] * X[i][2];
                    U[k][i][2] = 0.5 * m[4] * X[i][0] + 0.5 * m[3] * X[i][1] + m[2] * X[i][2];
                }
            }
        }
    }
    double F[NR][NP][B];
    for (int k = 0; k < NR; k++)
        for (int i = 0; i < NP; i++)
            for (int c = 0; c < B; c++) F[k][i][c] = 0.;
    const uint32_t *slots = a.eslot + (int64_t)e * NP * NP;
    for (int i = 0; i < NP; i++) {
        for (int j = 0; j <= i; j++) {
            double K[B * B];
            element_block<ET>(G, w, C, i, j, K);
            add_block64<B>(a.v64, a.nslots, slots[i * NP + j], K);
            double KT[B * B];
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++) KT[r * B + c] = K[c * B + r];
            if (j != i) add_block64<B>(a.v64, a.nslots, slots[j * NP + i], KT);
            if (macro) {
                for (int k = 0; k < NR; k++)
                    for (int r = 0; r < B; r++)
                        for (int c = 0; c < B; c++) {
                            F[k][i][r] -= K[r * B + c] * U[k][j][c];
                            if (j != i) F[k][j][r] -= KT[r * B + c] * U[k][i][c];
                        }
            }
        }
    }
    if (macro) {
        for (int i = 0; i < NP; i++) {
            const int64_t row = a.erow[(int64_t)e * NP + i];
            for (int k = 0; k < NR; k++)
                for (int c = 0; c < B; c++) a.b64[k * a.vplane + row * B + c] += F[k][i][c];
        }
    }
}

// ------------------------------------------------- gather assembly
// One thread per block slot: the slot's contributions, listed at batch build
// as codes (e*NPER + i)*NPER + j (element e, block row node i, column node j)
// in the reference's own order (solvers/assembly.cpp:657-735: elements in
// order, and inside one element (j,j), then (j,k), (k,j) for k > j), are
// recomputed and added with the same Kahan compensation, so each entry is
// rounded exactly as the serial loop rounds it.  Every slot is written once
// (padding slots get zeros): no read-modify-write, coalesced stores.
template <int ET> __device__ inline void load_element(const AsmArgs &a, int e, double (*X)[3], double *C)
{
    using E = Elem<ET>;
    constexpr int NP = E::NPER, S = E::S, D = E::DIM;
    const int p = a.eprob[e];
    for (int i = 0; i < NP; i++) {
        const int nd = a.econn[(int64_t)e * NP + i];
        for (int d = 0; d < D; d++) X[i][d] = a.coords[(int64_t)nd * D + d];
        if (D == 2) X[i][2] = 0.;
    }
    const double *Cp = a.cmat + (int64_t)(a.cmat_off[p] + a.ephase[e]) * S * S;
    const double f = stiffness_factor(a.edamage, e, a.frac_at, a.residual);
    for (int k = 0; k < S * S; k++) C[k] = a.edamage ? Cp[k] * f : Cp[k];
}

template <int ET>
__global__ __launch_bounds__(256) void k_assemble_gather(AsmArgs a, const uint32_t *cptr, const uint32_t *contrib)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int NP = E::NPER, B = E::B, S = E::S;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslots) return;
    double acc[B * B], comp[B * B];  // plane order: entry (r,c) at c*B + r
    for (int q = 0; q < B * B; q++) { acc[q] = 0.; comp[q] = 0.; }
    const uint32_t q1 = cptr[s + 1];
    for (uint32_t q = cptr[s]; q < q1; q++) {
        const uint32_t code = contrib[q];
        const int e = (int)(code / (NP * NP));
        const int i = (int)(code % (NP * NP)) / NP, j = (int)(code % NP);
        double X[NP][3], C[S * S];
        load_element<ET>(a, e, X, C);
        double G[E::NGP][6][3], w[E::NGP];
        element_grads<ET>(X, G, w);
        double ga[E::NGP][3], gb[E::NGP][3];
        pick_grads<ET>(G, i >= j ? i : j, ga);
        pick_grads<ET>(G, i >= j ? j : i, gb);
        double K[B * B];
        element_block_pair<ET>(ga, gb, w, C, K);
        for (int r = 0; r < B; r++)
            for (int c = 0; c < B; c++) {
                const double v = i >= j ? K[r * B + c] : K[c * B + r];
                double &x = acc[c * B + r], &cc = comp[c * B + r];
                const double y = v - cc;
                const double t = x + y;
                cc = (t - x) - y;
                x = t;
            }
    }
    double2 *p = reinterpret_cast<double2 *>(a.v64);
    constexpr int NPL = (B * B) / 2;
    for (int k = 0; k < NPL; k++) p[(int64_t)k * a.nslots + s] = make_double2(acc[2 * k], acc[2 * k + 1]);
    if constexpr ((B * B) % 2) reinterpret_cast<double *>(p + (int64_t)NPL * a.nslots)[s] = acc[B * B - 1];
}

// Macro-strain right-hand side, one thread per internal block row: the row's
// (element, local node) pairs in element order, f -= Ke(row, :) u_macro with
// the element row summed over its DOFs in order (the oracle's restatement of
// the homogenisation load).
template <int ET, int NR>
__global__ __launch_bounds__(256) void k_macro_rhs_gather(AsmArgs a, const uint32_t *rptr, const uint32_t *rlist,
                                                          int64_t nrows)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int NP = E::NPER, B = E::B, S = E::S, D = E::DIM;
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const uint32_t q0 = rptr[row], q1 = rptr[row + 1];
    if (q0 == q1) return;
    double f[NR][B];
    for (int k = 0; k < NR; k++)
        for (int r = 0; r < B; r++) f[k][r] = a.b64[k * a.vplane + row * B + r];
    for (uint32_t q = q0; q < q1; q++) {
        const uint32_t code = rlist[q];
        const int e = (int)(code / NP), i = (int)(code % NP);
        double X[NP][3], C[S * S];
        load_element<ET>(a, e, X, C);
        double G[E::NGP][6][3], w[E::NGP];
        element_grads<ET>(X, G, w);
        const int p = a.eprob[e];
        double U[NR][NP][B];
        for (int k = 0; k < NR; k++) {
            const double *m = a.macro + ((int64_t)p * NR + k) * S;
            for (int n = 0; n < NP; n++) {
                if constexpr (D == 2) {
                    U[k][n][0] = m[0] * X[n][0] + 0.5 * m[2] * X[n][1];
                    U[k][n][1] = 0.5 * m[2] * X[n][0] + m[1] * X[n][1];
                } else {
                    U[k][n][0] = m[0] * X[n][0] + 0.5 * m[5] * X[n][1] + 0.5 * m[4] * X[n][2];
                    U[k][n][1] = 0.5 * m[5] * X[n][0] + m[1] * X[n][1] + 0.5 * m[3] * X[n][2];
                    U[k][n][2] = 0.5 * m[4] * X[n][0] + 0.5 * m[3] * X[n][1] + m[2] * X[n][2];
                }
            }
        }
        double t[NR][B];
        for (int k = 0; k < NR; k++)
            for (int r = 0; r < B; r++) t[k][r] = 0.;
        double gi[E::NGP][3];
        pick_grads<ET>(G, i, gi);
#pragma unroll
        for (int j = 0; j < NP; j++) {
            double gj[E::NGP][3];
            for (int g = 0; g < E::NGP; g++)
                for (int d = 0; d < 3; d++) gj[g][d] = G[g][j][d];
            double K[B * B];
            if (i >= j) element_block_pair<ET>(gi, gj, w, C, K);
            else element_block_pair<ET>(gj, gi, w, C, K);
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++) {
                    const double v = i >= j ? K[r * B + c] : K[c * B + r];
                    for (int k = 0; k < NR; k++) t[k][r] = t[k][r] + v * U[k][j][c];
                }
        }
        for (int k = 0; k < NR; k++)
            for (int r = 0; r < B; r++) f[k][r] = f[k][r] - t[k][r];
    }
    for (int k = 0; k < NR; k++)
        for (int r = 0; r < B; r++) a.b64[k * a.vplane + row * B + r] = f[k][r];
}

// ------------------------------------------------- Dirichlet elimination
// solvers/assembly.cpp:125-253 restated per block row: for a free row i,
// b_i -= g_j A(i,j) and A(i,j) = 0 for every constrained column j (blocks in
// column order); a constrained row becomes a unit row with b_i = g_i.
template <int B, int NR>
__global__ __launch_bounds__(WG) void k_apply_bc(SolveArgs a, double *v64, double *b64, const uint8_t *bcmask,
                                                 const double *bcval, const int32_t *row_len, int ntiles)
{
    AMIE_NO_CONTRACT  // b -= g A as the reference rounds it: a product, then a subtraction
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (wave >= a.tile_nslices[t]) continue;
        const int s = a.tile_slice0[t] + wave;
        if (lane >= a.slice_rows[s]) continue;
        const int64_t row = (int64_t)s * WAVE + lane;
        bool rfix[B];
        bool any_r = false;
        for (int i = 0; i < B; i++) { rfix[i] = bcmask[row * B + i]; any_r |= rfix[i]; }
        // the row's own blocks only: the zero padding blocks of a SELL slice
        // must stay zero
        const int w = row_len[row];
        const int64_t sp = a.slice_ptr[s] + lane;
        for (int k = 0; k < w; k++) {
            const int64_t slot = sp + (int64_t)k * WAVE;
            const uint32_t c = a.col[slot];
            double blk[B * B];
            Blk<double, B>::load(v64, a.nslots, slot, blk);
            bool mod = false;
            for (int j = 0; j < B; j++) {
                const int64_t cd = (int64_t)c * B + j;
                if (!bcmask[cd]) continue;
                const double g = bcval[cd];
                for (int i = 0; i < B; i++) {
                    if (rfix[i]) continue;
                    for (int r = 0; r < NR; r++) b64[r * a.vplane + row * B + i] -= g * blk[j * B + i];
                    blk[j * B + i] = 0.;
                    mod = true;
                }
            }
            if (any_r) {
                for (int i = 0; i < B; i++) {
                    if (!rfix[i]) continue;
                    for (int j = 0; j < B; j++) blk[j * B + i] = ((int64_t)c * B + j == row * B + i) ? 1. : 0.;
                }
                mod = true;
            }
            if (mod) Blk<double, B>::store(v64, a.nslots, slot, blk);
        }
        for (int i = 0; i < B; i++)
            if (rfix[i])
                for (int r = 0; r < NR; r++) b64[r * a.vplane + row * B + i] = bcval[row * B + i];
    }
}

// S = diag(K)^-1/2 per dof (1 where the diagonal is not positive).
template <int B>
__global__ __launch_bounds__(WG) void k_diag_scale(SolveArgs a, const int64_t *dslot, double *S, int ntiles)
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (wave >= a.tile_nslices[t]) continue;
        const int s = a.tile_slice0[t] + wave;
        const int64_t row = (int64_t)s * WAVE + lane;
        if (lane >= a.slice_rows[s]) {
            for (int i = 0; i < B; i++) S[row * B + i] = 1.;
            continue;
        }
        double blk[B * B];
        Blk<double, B>::load(a.v64, a.nslots, dslot[row], blk);
        for (int i = 0; i < B; i++) {
            double d = blk[i * B + i];
            S[row * B + i] = d > 0. ? 1. / sqrt(d) : 1.;
        }
    }
}

// Scaled working matrix Ks = S K S in W, and the inverse of each scaled
// diagonal block (block Jacobi).
template <typename M, typename W, int B>
__global__ __launch_bounds__(WG) void k_scale_cast(SolveArgs a, void *vw, const int64_t *dslot, W *minv, int ntiles)
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (wave >= a.tile_nslices[t]) continue;
        const int s = a.tile_slice0[t] + wave;
        const int64_t row = (int64_t)s * WAVE + lane;
        const bool valid = lane < a.slice_rows[s];
        double sr[B];
        for (int i = 0; i < B; i++) sr[i] = a.S[row * B + i];
        const int w = a.slice_width[s];
        const int64_t sp = a.slice_ptr[s] + lane;
        for (int k = 0; k < w; k++) {
            const int64_t slot = sp + (int64_t)k * WAVE;
            const uint32_t c = a.col[slot];
            double blk[B * B];
            Blk<double, B>::load(a.v64, a.nslots, slot, blk);
            typedef typename Blk<M, B>::T T;
            T o[B * B];
            for (int j = 0; j < B; j++) {
                const double sc = a.S[(int64_t)c * B + j];
                for (int i = 0; i < B; i++) o[j * B + i] = valid ? (T)(sr[i] * blk[j * B + i] * sc) : (T)0;
            }
            if constexpr (std::is_same<M, Fix48>::value) {
                bool bad = false;
                for (int q = 0; q < B * B; q++) bad |= !(fabs((double)o[q]) < FIX48_LIMIT);
                if (bad) {
                    *a.fix48_bad = 1;   // every writer stores the same value
                    for (int q = 0; q < B * B; q++) o[q] = 0.;
                }
            }
            if (vw) Blk<M, B>::store(vw, a.nslots, slot, o);
        }
        if (!valid || !minv) continue;
        double d[B * B];
        Blk<double, B>::load(a.v64, a.nslots, dslot[row], d);
        double m[B * B]; // row-major scaled diagonal block
        for (int i = 0; i < B; i++)
            for (int j = 0; j < B; j++) m[i * B + j] = sr[i] * d[j * B + i] * sr[j];
        double inv[B * B];
        if constexpr (B == 2) {
            double dt = 1. / (m[0] * m[3] - m[1] * m[2]);
            inv[0] = m[3] * dt; inv[1] = -m[1] * dt; inv[2] = -m[2] * dt; inv[3] = m[0] * dt;
        } else {
            double c0 = m[4] * m[8] - m[5] * m[7], c1 = m[2] * m[7] - m[1] * m[8], c2 = m[1] * m[5] - m[2] * m[4];
            double c3 = m[5] * m[6] - m[3] * m[8], c4 = m[0] * m[8] - m[2] * m[6], c5 = m[2] * m[3] - m[0] * m[5];
            double c6 = m[3] * m[7] - m[4] * m[6], c7 = m[1] * m[6] - m[0] * m[7], c8 = m[0] * m[4] - m[1] * m[3];
            double dt = 1. / (m[0] * c0 + m[1] * c3 + m[2] * c6);
            inv[0] = c0 * dt; inv[1] = c1 * dt; inv[2] = c2 * dt; inv[3] = c3 * dt; inv[4] = c4 * dt;
            inv[5] = c5 * dt; inv[6] = c6 * dt; inv[7] = c7 * dt; inv[8] = c8 * dt;
        }
        for (int q = 0; q < B * B; q++) minv[row * B * B + q] = (W)inv[q];
    }
}

// ------------------------------------------------------ strain / damage
struct StrainArgs {
    const double *coords;
    const int32_t *econn;
    const int32_t *erow;
    const int32_t *eprob;
    const double *macro;      // nullable
    const double *x64;        // solution planes (vplane entries each)
    int64_t vplane;
    int nrhs;
};

// Centroid strain, elements/integrable_entity.cpp:977-1097 (Voigt,
// engineering shear), displacement = macro part + solved part.
template <int ET>
__device__ inline void element_strain_at(const StrainArgs &a, int e, int k, double px, double py, double *eps)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int NP = E::NPER, D = E::DIM, B = E::B;
    double X[NP][3];
    for (int i = 0; i < NP; i++) {
        const int nd = a.econn[(int64_t)e * NP + i];
        for (int d = 0; d < D; d++) X[i][d] = a.coords[(int64_t)nd * D + d];
        if (D == 2) X[i][2] = 0.;
    }
    double m[9];
    centroid_jinv<ET>(X, m);   // the element's inverse Jacobian cache is made at the centroid
    double dxi[6], deta[6], dzeta[6];
    dshape<ET>(px, py, dxi, deta, dzeta);
    double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}; // [component][reference direction]
    for (int j = 0; j < NP; j++) {
        const int64_t row = a.erow[(int64_t)e * NP + j];
        for (int c = 0; c < B; c++) {
            double u = a.x64[k * a.vplane + row * B + c];
            g[c][0] += dxi[j] * u; g[c][1] += deta[j] * u;
            if (D == 3) g[c][2] += dzeta[j] * u;
        }
    }
    const double *mac = a.macro ? a.macro + ((int64_t)a.eprob[e] * a.nrhs + k) * E::S : nullptr;
    if constexpr (D == 2) {
        eps[0] = g[0][0] * m[0] + g[0][1] * m[1];
        eps[1] = g[1][0] * m[2] + g[1][1] * m[3];
        eps[2] = g[0][0] * m[2] + g[0][1] * m[3] + g[1][0] * m[0] + g[1][1] * m[1];
        if (mac) for (int q = 0; q < 3; q++) eps[q] += mac[q];
    } else {
        auto d = [&](int c, int r) { return g[c][0] * m[r * 3] + g[c][1] * m[r * 3 + 1] + g[c][2] * m[r * 3 + 2]; };
        eps[0] = d(0, 0); eps[1] = d(1, 1); eps[2] = d(2, 2);
        eps[3] = d(1, 2) + d(2, 1); eps[4] = d(0, 2) + d(2, 0); eps[5] = d(1, 0) + d(0, 1);
        if (mac) for (int q = 0; q < 6; q++) eps[q] += mac[q];
    }
}

template <int ET> __device__ inline void element_strain(const StrainArgs &a, int e, int k, double *eps)
{
    if constexpr (Elem<ET>::DIM == 2) element_strain_at<ET>(a, e, k, 1. / 3., 1. / 3., eps);
    else element_strain_at<ET>(a, e, k, .25, .25, eps);
}

// Strain and stress at every Gauss point of each element's rule
// (ElementState::getField TOTAL_STRAIN_FIELD / REAL_STRESS_FIELD at the
// Gauss points, elements/integrable_entity.cpp:977-1097,1379-1394; the
// stress is strain . C (1 - d), or C * residual once fractured).
template <int ET>
__global__ __launch_bounds__(256) void k_gauss_fields(StrainArgs a, int e0, int n, int k, const int32_t *ephase,
                                                      const double *edamage, const int32_t *cmat_off,
                                                      const double *cmat, double frac_at, double residual,
                                                      double *strain, double *stress)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int S = E::S, NG = E::NGP;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = e0 + i;
    const double *Cp = cmat + (int64_t)(cmat_off[a.eprob[e]] + ephase[e]) * S * S;
    const double f = stiffness_factor(edamage, e, frac_at, residual);
    for (int g = 0; g < NG; g++) {
        double gx, gy, gw;
        gauss<ET>(g, gx, gy, gw);
        double eps[S];
        element_strain_at<ET>(a, e, k, gx, gy, eps);
        double *o = strain + ((int64_t)i * NG + g) * S;
        for (int q = 0; q < S; q++) o[q] = eps[q];
        if (stress) {
            double *t = stress + ((int64_t)i * NG + g) * S;
            for (int j = 0; j < S; j++) {
                double acc = 0.;
                for (int q = 0; q < S; q++) acc += eps[q] * (Cp[q * S + j] * f);
                t[j] = acc;
            }
        }
    }
}

template <int ET>
__global__ __launch_bounds__(256) void k_strain(StrainArgs a, int e0, int n, int k, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    element_strain<ET>(a, e0 + i, k, out + (int64_t)i * Elem<ET>::S);
}

// Principal strains, Mazars equivalent strain (tension branch,
// physics/fracturecriteria/mazars.cpp:124), kappa history and the damage law
// d = 1 - (k0/kappa) exp(-(kappa - k0)/kf) for kappa > k0, combined with the
// initial damage as max(d0, d_law), never decreasing, capped at dmax.
__device__ inline double mazars_equivalent(int dim, const double *s)
{
    double ev[3];
    if (dim == 2) {
        double a = s[0], b = s[1], c = 0.5 * s[2];
        double mm = 0.5 * (a + b), r = sqrt(0.25 * (a - b) * (a - b) + c * c);
        ev[0] = mm + r; ev[1] = mm - r; ev[2] = 0.;
    } else {
        double a11 = s[0], a22 = s[1], a33 = s[2], a23 = 0.5 * s[3], a13 = 0.5 * s[4], a12 = 0.5 * s[5];
        double p1 = a12 * a12 + a13 * a13 + a23 * a23;
        double q = (a11 + a22 + a33) / 3.;
        if (p1 <= 1e-300) { ev[0] = a11; ev[1] = a22; ev[2] = a33; }
        else {
            double p2 = (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + (a33 - q) * (a33 - q) + 2. * p1;
            double p = sqrt(p2 / 6.);
            double b11 = (a11 - q) / p, b22 = (a22 - q) / p, b33 = (a33 - q) / p;
            double b12 = a12 / p, b13 = a13 / p, b23 = a23 / p;
            double detB = b11 * (b22 * b33 - b23 * b23) - b12 * (b12 * b33 - b23 * b13) + b13 * (b12 * b23 - b22 * b13);
            double rr = fmin(1., fmax(-1., detB / 2.));
            double phi = acos(rr) / 3.;
            ev[0] = q + 2. * p * cos(phi);
            ev[2] = q + 2. * p * cos(phi + 2. * 3.14159265358979323846 / 3.);
            ev[1] = 3. * q - ev[0] - ev[2];
        }
    }
    double eq = 0.;
    for (int i = 0; i < 3; i++) { double t = fmax(0., ev[i]); eq += t * t; }
    return sqrt(eq);
}

template <int ET>
__global__ __launch_bounds__(256) void k_damage(StrainArgs a, int n, const double *d0, double *kappa, double *d,
                                               double k0, double kf, double dmax, unsigned long long *grown,
                                               const double *eps_in)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double eps[6];
    // eps_in: the non-local strains (else the element's own)
    if (eps_in)
        for (int q = 0; q < Elem<ET>::S; q++) eps[q] = eps_in[(int64_t)e * Elem<ET>::S + q];
    else
        element_strain<ET>(a, e, 0, eps);
    double eq = mazars_equivalent(Elem<ET>::DIM, eps);
    double kap = fmax(kappa[e], eq);
    kappa[e] = kap;
    double dl = kap > k0 ? 1. - (k0 / kap) * exp(-(kap - k0) / kf) : 0.;
    double dn = fmin(dmax, fmax(d0 ? d0[e] : 0., fmax(d[e], dl)));
    if (dn > d[e]) atomicAdd(grown, 1ull);
    d[e] = dn;
}

// Non-local (smoothed) strains of linear elements (P1, Tet4), restating
// Mesh::getSmoothedFields (mesher/mesh.h:1384-1486, non-space-time branch)
// over the neighbourhoods of Mesh::generateCache (mesher/mesh.h:546-640) set
// up by FractureCriterion::initialiseCache
// (physics/fracturecriteria/fracturecriterion.cpp:171-190): element e
// averages the strains of the same-phase, unfractured elements whose
// centroid lies within rho_e = 2.5 R_e + 1.5 r_e of its centroid (R_e =
// max(radius, 1.1 r_e), r_e its circumradius), each counted with its measure
// times its cache weight, the mean over its vertices of
// max(w(|v - c_e|^2 / R_e^2) * measure, 0) with the smoothing kernel w
// (fracturecriterion.cpp:43-61).  The candidates come
// from a uniform grid of element centroids per problem (cell = the largest
// rho of the problem).
struct NonlocalArgs {
    const double *geo;         // per element: cx, cy, cz, R, rho, measure, r, rho2
    const int32_t *econn;
    const double *coords;
    const int32_t *ephase;
    const double *edamage;     // nullable
    double threshold;          // fractured: damage >= threshold
    const int32_t *eprob;
    const double *grid;        // per problem: origin x, y, z, cell size
    const int32_t *gdim;       // per problem: cells along x, y, z, first cell
    const int32_t *cell_start; // per cell (all problems): first entry in cell_elem; one more at the end
    const int32_t *cell_elem;  // element ids, cell by cell
    int kernel;                // 0 QUARTIC_COMPACT, 1 GAUSSIAN_NONCOMPACT
};
constexpr int GEO_STRIDE = 8;

__device__ inline double smoothing_weight(int kernel, double rrn)
{
    if (kernel == 1) return exp(-rrn);
    const double t = rrn / 2.;
    return 1. - t >= 0. ? (t - 1.) * (t - 1.) : 0.;
}

// LANES = 1: a thread per element; LANES = 64: a wavefront per element, its
// lanes taking every 64th candidate of each cell and the partial sums
// added across the wavefront at the end (for neighbourhoods of hundreds of
// elements, where a thread per element leaves the GPU idle).
template <int ET, int LANES>
__global__ __launch_bounds__(256) void k_smooth(NonlocalArgs a, const double *strain, int n, double *out)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int S = E::S, NP = E::NPER, D = E::DIM;
    const int64_t e = LANES == 1 ? (int64_t)blockIdx.x * blockDim.x + threadIdx.x
                                 : (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = LANES == 1 ? 0 : (int)(threadIdx.x % WAVE);
    if (e >= n) return;   // uniform over a wavefront when LANES = 64
    const double *ge = a.geo + e * GEO_STRIDE;
    const double c[3] = {ge[0], ge[1], ge[2]};
    const double R = ge[3], rho = ge[4];
    const int p = a.eprob[e];
    const int ph = a.ephase ? a.ephase[e] : 0;
    const double *g = a.grid + (int64_t)p * 4;
    const int *gd = a.gdim + (int64_t)p * 4;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
        lo[d] = max(0, (int)floor((c[d] - rho - g[d]) / g[3]));
        hi[d] = min(gd[d] - 1, (int)floor((c[d] + rho - g[d]) / g[3]));
    }
    double acc[S], cor[S], sum = 0.;
    for (int q = 0; q < S; q++) acc[q] = cor[q] = 0.;
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++) {
                const int cell = gd[3] + (z * gd[1] + y) * gd[0] + x;
                for (int j = a.cell_start[cell] + lane; j < a.cell_start[cell + 1]; j += LANES) {
                    const int i = a.cell_elem[j];
                    if (a.ephase && a.ephase[i] != ph) continue;
                    if (a.edamage && a.edamage[i] >= a.threshold) continue;
                    const double *gi = a.geo + (int64_t)i * GEO_STRIDE;
                    double d2 = 0.;
                    for (int d = 0; d < D; d++) d2 += (gi[d] - c[d]) * (gi[d] - c[d]);
                    // Circle::in strict (geometry_2D.cpp:1472-1484); Sphere::in
                    // with its 1e-12 tolerance (geometry_3D.cpp:1551-1566)
                    const bool in = D == 2 ? d2 < rho * rho : d2 < rho * rho + 2. * rho * 1e-12 + 1e-24;
                    if (!in) continue;
                    double coef = 0.;
                    for (int v = 0; v < NP; v++) {
                        const int nd = a.econn[(int64_t)i * NP + v];
                        double r2 = 0.;
                        for (int d = 0; d < D; d++) {
                            const double t = a.coords[(int64_t)nd * D + d] - c[d];
                            r2 += t * t;
                        }
                        coef += fmax(smoothing_weight(a.kernel, r2 / (R * R)) * gi[5], 0.);
                    }
                    coef /= NP;
                    // ElementState::getAverageField returns measure * weight
                    // (elements/integrable_entity.cpp:2026-2133), which
                    // Mesh::getSmoothedFields Kahan-adds (mesher/mesh.h:1453-1462)
                    const double v = gi[5] * coef;
                    for (int q = 0; q < S; q++) {
                        const double yk = strain[(int64_t)i * S + q] * v - cor[q];
                        const double t = acc[q] + yk;
                        cor[q] = (t - acc[q]) - yk;
                        acc[q] = t;
                    }
                    sum += v;
                }
            }
    if constexpr (LANES > 1) {
        for (int q = 0; q < S; q++) acc[q] = wave_sum(acc[q] - cor[q]);
        sum = wave_sum(sum);
        if (lane) return;
    }
    for (int q = 0; q < S; q++) out[e * S + q] = sum > 1e-12 ? acc[q] / sum : 0.;
}

// ------------------------------------------------- event-driven damage
// FeatureTree's damage iteration for StiffnessAndFracture + NonLocalMazars +
// FiberBasedIsotropicLinearDamage (features/features.cpp:4875-5290,
// 6150-6373): after every solve each element of a criterion phase is scored
// on its smoothed strain, then one pass visits the elements and increments
// the damage of those at their neighbourhood's maximum score.
struct EventArgs {
    const int32_t *crit;      // per material: 1 = the criterion and damage model above
    const double *prm;        // per material: k0, E, nu, Gf
    const int32_t *cmat_off;  // per problem: first C matrix
    const double *cmat;
    const int32_t *prob_e0, *prob_ne;   // per problem: first element, elements
    double frac_at;           // fractured at d >= frac_at (0.6)
    double fraction;          // damage added per event (0.1)
    double tol;               // FractureCriterion::scoreTolerance (2e-3)
    int64_t first_deferred;   // the pass visits positions >= first_deferred first
};

// toPrincipal (elements/integrable_entity.cpp:475-560): doubled = 1 for an
// engineering strain (shear entries 2 e_ij).
template <int D> __device__ inline void principal_values(const double *v, bool doubled, double *out)
{
    AMIE_NO_CONTRACT
    if constexpr (D == 2) {
        const double tr = v[0] + v[1];
        const double det = doubled ? v[0] * v[1] - 0.25 * v[2] * v[2] : v[0] * v[1] - v[2] * v[2];
        const double delta = sqrt(tr * tr - 4. * det);
        const double angle = doubled ? 0.5 * atan2(0.5 * v[2], v[0] - v[1]) : 0.5 * atan2(v[2], v[0] - v[1]);
        if (cos(angle) < 0) {
            out[0] = (tr + delta) * .5;
            out[1] = (tr - delta) * .5;
        } else {
            out[0] = (tr - delta) * .5;
            out[1] = (tr + delta) * .5;
        }
    } else {
        const double h = doubled ? 0.5 : 1.;
        const double m00 = v[0], m11 = v[1], m22 = v[2], m12 = v[3] * h, m02 = v[4] * h, m01 = v[5] * h;
        const double trmat = -(m00 + m11 + m22);
        const double detmat = -2.0 * m01 * m02 * m12 + m00 * m12 * m12 + m11 * m02 * m02 + m22 * m01 * m01 - m00 * m11 * m22;
        const double m2mat = (m00 * m11 + m11 * m22 + m22 * m00) - m02 * m02 - m01 * m01 - m12 * m12;
        const double q = m2mat / 3. - trmat * trmat / 9.;
        const double r = (trmat * m2mat - 3. * detmat) / 6. - trmat * trmat * trmat / 27.;
        const double d = q * q * q + r * r;
        const double r0 = pow(r * r - d, 1. / 6.);
        double phi = atan2(sqrt(-1. * d), r) / 3.;
        if (fabs(phi) < 1e-12) phi = 0.;
        if (phi < 0.) phi += M_PI;
        const double som = r0 * cos(phi), dif = r0 * sin(phi);
        out[0] = 2. * som - trmat / 3.;
        out[1] = -som - trmat / 3. - dif * sqrt(3.);
        out[2] = -som - trmat / 3. + dif * sqrt(3.);
    }
}

__device__ inline double pos_part(double x) { return 0.5 * (fabs(x) + x); }

// NonLocalMazars::grade (physics/fracturecriteria/mazars.cpp:73-151) of the
// criterion copy every element holds (compressive parameters -1, so in 3D
// not tension-only; the plane-stress branch is), FractureCriterion::step
// (fracturecriterion.cpp:534-560): score = pseudo-damage - d, -1 when
// fractured.  sbar: the smoothed strains; elements of a phase without a
// criterion get 0.
template <int ET>
__global__ __launch_bounds__(256) void k_mazars_score(EventArgs ev, const int32_t *ephase, const int32_t *eprob,
                                                      const double *edamage, const double *sbar, int n, double *score)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int S = E::S, D = E::DIM;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int ph = ephase[e];
    if (!ev.crit[ph]) { score[e] = 0.; return; }
    const double d = edamage[e];
    if (d >= ev.frac_at) { score[e] = -1.; return; }
    const double k0 = fabs(ev.prm[4 * ph]), Ey = ev.prm[4 * ph + 1], nu = ev.prm[4 * ph + 2], Gf = ev.prm[4 * ph + 3];
    double eps[S];
    for (int q = 0; q < S; q++) eps[q] = sbar[(int64_t)e * S + q];
    const double *Cp = ev.cmat + (int64_t)(ev.cmat_off[eprob[e]] + ph) * S * S;
    double sig[S], ps[3] = {0., 0., 0.}, pe[3] = {0., 0., 0.};
    for (int j = 0; j < S; j++) {
        double t = 0.;
        for (int i = 0; i < S; i++) t += eps[i] * (Cp[i * S + j] * (1. - d));
        sig[j] = t;
    }
    principal_values<D>(sig, false, ps);
    principal_values<D>(eps, true, pe);
    double trt = 0., trc = 0., sc2 = 0.;
    for (int i = 0; i < D; i++) {
        trt += 0.5 * (fabs(ps[i]) + ps[i]);
        trc += fmin(0.0, ps[i]);
        sc2 += fmin(0.0, ps[i]) * fmin(0.0, ps[i]);
    }
    double gamma = 1.0;
    if (trc < -1e-12 && fabs(trt / trc) < 1e-8) gamma = sqrt(sc2) / (-trc);
    double kappa, talpha;
    if constexpr (D == 2) {
        const double ezz = -nu * (pe[0] + pe[1]) / (1. - nu);
        const double a0 = pos_part(pe[0]), a1 = pos_part(pe[1]), a2 = pos_part(ezz);
        const double m = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
        kappa = gamma * (0.0 < m ? m : 0.0);
        talpha = 1.0;
    } else {
        const double ed = Ey * (1 - d);
        const double q0 = (pos_part(ps[0]) - nu * (pos_part(ps[1]) + pos_part(ps[2]))) / ed;
        const double q1 = (pos_part(ps[1]) - nu * (pos_part(ps[0]) + pos_part(ps[2]))) / ed;
        const double q2 = (pos_part(ps[2]) - nu * (pos_part(ps[0]) + pos_part(ps[1]))) / ed;
        const double a0 = pos_part(pe[0]), a1 = pos_part(pe[1]), a2 = pos_part(pe[2]);
        const double m = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
        kappa = gamma * (0. < m ? m : 0.);
        talpha = (q0 * a0 + q1 * a1 + q2 * a2) / (kappa * kappa);
        if (talpha < 1.0e-3) talpha = 0.;
    }
    const double calpha = 1.0 - talpha;
    const double cstress = -1., cstrain = -1.;
    const double Bt = (0.1 * Ey * k0) / (Gf - Ey * 0.1 * k0 * k0 * 0.5);
    const double Bc = -1. / (sqrt(2.) * nu * cstrain);
    const double Ac = -(Ey * k0 + sqrt(2.) * cstress * nu) / ((Ey * exp(Bc * k0 - 1.) / Bc) - Ey * k0);
    const double kt = k0 < kappa ? kappa : k0;
    const double ps_ = talpha * (1. - (k0 / kt) * (exp(-Bt * (kt - k0)))) +
                       calpha * (1. - (k0 * (1 - Ac) / kt) - Ac * (exp(-Bc * (kt - k0))));
    const double pseudo = 1.0 < ps_ ? 1.0 : ps_;
    score[e] = pseudo - d;
}

// position of element e in the order the reference's pass visits its
// problem's elements: positions >= fd first, then 0 .. fd-1
__device__ inline int64_t visit_rank(int64_t local, int64_t ne, int64_t fd)
{
    const int64_t f = fd < 0 ? 0 : (fd > ne ? ne : fd);
    return local >= f ? local - f : local + (ne - f);
}

// One round of the pass's decisions (FiberBasedIsotropicLinearDamage::step,
// physics/damagemodels/fiberbasedisotropiclineardamage.cpp:58-92, with
// FractureCriterion::getMaxScoreInNeighbourhood, fracturecriterion.cpp:
// 251-280): element e takes an event when it is in a criterion phase, not
// fractured, scores above 0 and lies within tol of the largest score among
// itself and the same-phase neighbours within rho2_e (the second cache of
// FractureCriterion::initialiseCache, :171-190) that are not fractured when
// the pass reaches e: fractured before the pass (dprev >= frac_at), or made
// fractured by this pass at an element visited earlier (newfrac, the
// previous round's guess).  Rounds repeat until newfrac stops growing; the
// decisions then equal those of the reference's serial pass.
template <int ET, int LANES>
__global__ __launch_bounds__(256) void k_fiber_decide(NonlocalArgs a, EventArgs ev, const double *score,
                                                      const double *dprev, const uint8_t *newfrac, int n, uint8_t *inc)
{
    using E = Elem<ET>;
    constexpr int D = E::DIM;
    const int64_t e = LANES == 1 ? (int64_t)blockIdx.x * blockDim.x + threadIdx.x
                                 : (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = LANES == 1 ? 0 : (int)(threadIdx.x % WAVE);
    if (e >= n) return;   // uniform over a wavefront when LANES = 64
    const int ph = a.ephase[e];
    const double s = score[e];
    if (!ev.crit[ph] || !(s > 0.) || dprev[e] >= ev.frac_at) {
        if (!lane) inc[e] = 0;
        return;
    }
    const int p = a.eprob[e];
    const int64_t e0 = ev.prob_e0[p], ne = ev.prob_ne[p];
    const int64_t re = visit_rank(e - e0, ne, ev.first_deferred);
    const double *ge = a.geo + e * GEO_STRIDE;
    const double c[3] = {ge[0], ge[1], ge[2]};
    const double rho = ge[7];
    const double *g = a.grid + (int64_t)p * 4;
    const int *gd = a.gdim + (int64_t)p * 4;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
        lo[d] = max(0, (int)floor((c[d] - rho - g[d]) / g[3]));
        hi[d] = min(gd[d] - 1, (int)floor((c[d] + rho - g[d]) / g[3]));
    }
    double mx = s;
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++) {
                const int cell = gd[3] + (z * gd[1] + y) * gd[0] + x;
                for (int j = a.cell_start[cell] + lane; j < a.cell_start[cell + 1]; j += LANES) {
                    const int i = a.cell_elem[j];
                    if (a.ephase[i] != ph) continue;
                    const double *gi = a.geo + (int64_t)i * GEO_STRIDE;
                    double d2 = 0.;
                    for (int d = 0; d < D; d++) d2 += (gi[d] - c[d]) * (gi[d] - c[d]);
                    const bool in = D == 2 ? d2 < rho * rho : d2 < rho * rho + 2. * rho * 1e-12 + 1e-24;
                    if (!in || dprev[i] >= ev.frac_at) continue;
                    if (newfrac[i] && visit_rank(i - e0, ne, ev.first_deferred) < re) continue;
                    const double si = score[i];
                    mx = mx < si ? si : mx;   // std::max(mx, si): a NaN score is passed over
                }
            }
    if constexpr (LANES > 1) {
        // every lane started from s > 0, so no partial maximum is NaN
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(mx, o, WAVE);
            mx = mx < t ? t : mx;
        }
        if (lane) return;
    }
    inc[e] = fabs(s - mx) < ev.tol ? 1 : 0;
}

// The elements an event fractures (d + fraction, capped at 1, reaching
// frac_at); count = their number.
__global__ __launch_bounds__(256) void k_fiber_mark(EventArgs ev, const double *dprev, const uint8_t *inc, int n,
                                                    uint8_t *newfrac, unsigned long long *count)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const bool f = inc[e] && fmin(dprev[e] + ev.fraction, 1.) >= ev.frac_at;
    newfrac[e] = f ? 1 : 0;
    if (f) atomicAdd(count, 1ull);
}

// The pass's events: d += fraction, capped at 1; count = their number.
__global__ __launch_bounds__(256) void k_fiber_commit(EventArgs ev, const uint8_t *inc, int n, double *d,
                                                      unsigned long long *count)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    if (!inc[e]) return;
    d[e] = fmin(d[e] + ev.fraction, 1.);
    atomicAdd(count, 1ull);
}

// --- IsotropicLinearDamage, Newton branch (physics/damagemodels/
// isotropiclineardamage.cpp:64-92, postProcess :150-162) -----------------
//
// Scores are ordered as unsigned keys (sign bit flipped for positives, all
// bits for negatives) so that an integer atomicMax gives the double maximum.
__device__ inline unsigned long long score_key(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : u | (1ull << 63);
}
__device__ inline double key_score(unsigned long long k)
{
    return __longlong_as_double((long long)((k >> 63) ? k & ~(1ull << 63) : ~k));
}

// tmax[p] = the largest score over problem p's criterion elements, at least
// -1 (FeatureTree's maxScoreInit before the pass); tmax starts at key(-1).
__global__ __launch_bounds__(256) void k_score_max(EventArgs ev, const int32_t *ephase, const int32_t *eprob,
                                                   const double *score, int n, unsigned long long *tmax)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || !ev.crit[ephase[e]]) return;
    atomicMax(tmax + eprob[e], score_key(score[e]));
}

// One element of the pass.  T = tmax of its problem; e is in the damaging
// window when T > 0, s > 0 and T - s <= 4 tol.  Its neighbourhood (the
// second cache, as in k_fiber_decide) splits into the window and the rest;
// m = the largest score of the rest (max(0, T) when the rest is empty) and
// delta = max(s - m, T - s).  Unless |T| < tol/2 and (e is out of the window
// or |s| < tol/20), an element of the window with s > tol/20 is changed by
// clamp(min(delta, s (1 - d)), 0, 1) (inc[e], flag[e] = 1).  No decision
// reads damage written in the pass: the increments are added afterwards.
template <int ET, int LANES>
__global__ __launch_bounds__(256) void k_newton_decide(NonlocalArgs a, EventArgs ev, const double *score,
                                                       const double *d, const unsigned long long *tmax, int n,
                                                       double *inc, uint8_t *flag, unsigned long long *count)
{
    AMIE_NO_CONTRACT
    using E = Elem<ET>;
    constexpr int D = E::DIM;
    const int64_t e = LANES == 1 ? (int64_t)blockIdx.x * blockDim.x + threadIdx.x
                                 : (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = LANES == 1 ? 0 : (int)(threadIdx.x % WAVE);
    if (e >= n) return;   // uniform over a wavefront when LANES = 64
    const int ph = a.ephase[e];
    if (!ev.crit[ph]) {
        if (!lane) flag[e] = 0;
        return;
    }
    const int p = a.eprob[e];
    const double T = key_score(tmax[p]);
    const double s = score[e];
    const double win = 4. * ev.tol;
    const bool inset = T > 0. && T - s <= win && s > 0.;
    double delta = 0.;
    if (inset) {
        const double *ge = a.geo + e * GEO_STRIDE;
        const double c[3] = {ge[0], ge[1], ge[2]};
        const double rho = ge[7];
        const double *g = a.grid + (int64_t)p * 4;
        const int *gd = a.gdim + (int64_t)p * 4;
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int k = 0; k < D; k++) {
            lo[k] = max(0, (int)floor((c[k] - rho - g[k]) / g[3]));
            hi[k] = min(gd[k] - 1, (int)floor((c[k] + rho - g[k]) / g[3]));
        }
        int found = 0;
        double m = 0.;
        for (int z = lo[2]; z <= hi[2]; z++)
            for (int y = lo[1]; y <= hi[1]; y++)
                for (int x = lo[0]; x <= hi[0]; x++) {
                    const int cell = gd[3] + (z * gd[1] + y) * gd[0] + x;
                    for (int j = a.cell_start[cell] + lane; j < a.cell_start[cell + 1]; j += LANES) {
                        const int i = a.cell_elem[j];
                        if (a.ephase[i] != ph) continue;
                        const double *gi = a.geo + (int64_t)i * GEO_STRIDE;
                        double d2 = 0.;
                        for (int k = 0; k < D; k++) d2 += (gi[k] - c[k]) * (gi[k] - c[k]);
                        const bool in = D == 2 ? d2 < rho * rho : d2 < rho * rho + 2. * rho * 1e-12 + 1e-24;
                        if (!in) continue;
                        const double si = score[i];
                        if (T - si <= win && si > 0.) continue;   // in the damaging window
                        m = found ? (si < m ? m : si) : si;
                        found = 1;
                    }
                }
        if constexpr (LANES > 1) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double tm = __shfl_xor(m, o, WAVE);
                const int tf = __shfl_xor(found, o, WAVE);
                if (tf) m = found ? (tm < m ? m : tm) : tm;
                found |= tf;
            }
        }
        if (!found) m = 0. < T ? T : 0.;
        delta = s - m < T - s ? T - s : s - m;
    }
    if (lane) return;
    bool ch = false;
    if (!(fabs(T) < .5 * ev.tol && ((inset && fabs(s) < .05 * ev.tol) || !inset)) && s > .05 * ev.tol && inset) {
        const double cap = s * (1. - d[e]);
        const double v = cap < delta ? cap : delta;
        inc[e] = v < 0. ? 0. : (1. < v ? 1. : v);
        ch = true;
        atomicAdd(count, 1ull);
    }
    flag[e] = ch ? 1 : 0;
}

// postProcess: d += inc, capped at 1, for the elements the pass changed.
__global__ __launch_bounds__(256) void k_newton_commit(const uint8_t *flag, const double *inc, int n, double *d)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || !flag[e]) return;
    const double v = d[e] + inc[e];
    d[e] = 1. < v ? 1. : v;
}

// Volume average of C_e (1 - d_e) eps_e per problem and right-hand side:
// one workgroup per problem, fixed summation order.
template <int ET>
__global__ __launch_bounds__(256) void k_average_stress(StrainArgs a, const int32_t *prob_e0, const int32_t *prob_ne,
                                                        const int32_t *ephase, const double *edamage,
                                                        const int32_t *cmat_off, const double *cmat, double *out,
                                                        double frac_at, double residual)
{
    using E = Elem<ET>;
    constexpr int S = E::S, NP = E::NPER, D = E::DIM;
    __shared__ double lds[4][MAXQ];
    const int p = blockIdx.x;
    const int e0 = prob_e0[p], ne = prob_ne[p];
    for (int k = 0; k < a.nrhs; k++) {
        double acc[S + 1];
        for (int q = 0; q <= S; q++) acc[q] = 0.;
        for (int i = threadIdx.x; i < ne; i += blockDim.x) {
            const int e = e0 + i;
            double eps[6];
            element_strain<ET>(a, e, k, eps);
            double X[NP][3];
            for (int j = 0; j < NP; j++) {
                const int nd = a.econn[(int64_t)e * NP + j];
                for (int d = 0; d < D; d++) X[j][d] = a.coords[(int64_t)nd * D + d];
            }
            double meas;
            if constexpr (D == 2) {
                const int v1 = NP == 6 ? 2 : 1, v2 = NP == 6 ? 4 : 2;
                meas = 0.5 * fabs((X[v1][0] - X[0][0]) * (X[v2][1] - X[0][1]) - (X[v1][1] - X[0][1]) * (X[v2][0] - X[0][0]));
            } else {
                double u[3], v[3], w[3];
                for (int d = 0; d < 3; d++) { u[d] = X[1][d] - X[0][d]; v[d] = X[2][d] - X[0][d]; w[d] = X[3][d] - X[0][d]; }
                meas = fabs(u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0])) / 6.;
            }
            const double *C = cmat + (int64_t)(cmat_off[p] + ephase[e]) * S * S;
            const double f = stiffness_factor(edamage, e, frac_at, residual);
            for (int r = 0; r < S; r++) {
                double t = 0.;
                for (int c = 0; c < S; c++) t += C[r * S + c] * eps[c];
                acc[r] += meas * f * t;
            }
            acc[S] += meas;
        }
        wg_sum_all<S + 1>(acc, lds);
        if (threadIdx.x == 0)
            for (int r = 0; r < S; r++) out[((int64_t)p * a.nrhs + k) * S + r] = acc[r] / acc[S];
        __syncthreads();
    }
}

} // namespace amie
