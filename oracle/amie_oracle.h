/* CPU oracle for the AMIE FE-solve hot path: a plain C++ (FP64, OpenMP)
 * restatement of the reference algorithm, used ONLY as the checker by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing on the
 * product path links or calls it.
 *
 * Layout conventions (shared with the product C ABI, include/amie_gpu.h):
 *  - element types: 1 = P1 triangle (3 nodes), 2 = P2 triangle (6 nodes,
 *    [v0, m01, v1, m12, v2, m20]), 3 = linear tetrahedron (4 nodes);
 *  - a node-block CSR in the reference layout: row_size[nrows],
 *    column_index[nnzb] (block column = node id), array[nnzb*b*colLen] with
 *    each b x b block column-major and columns padded to colLen = b + b%2
 *    (sparse/sparse_matrix.cpp:59-66, sparse/sparse_vector.cpp:27-58);
 *  - DOF = node*b + axis.
 */
#ifndef AMIE_ORACLE_H
#define AMIE_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* OpenMP threads of the oracle (the reference's Kahan dot products are
 * reduced per thread, so bit-level agreement with golden vectors made with
 * one thread needs one thread here). */
void or_set_threads(int n);
int or_get_threads(void);

/* utilities/tensor.cpp:1845-1931 (YOUNG_POISSON); dim 2 -> 3x3, 3 -> 6x6, row-major. */
void or_cauchy_green(double E, double nu, int dim, int plane_strain, double *C);

/* geometry/geometry_2D.cpp:476-497 and geometry/geometry_3D.cpp:254-268: local
 * node order after the reference's orientation rule.  perm[k] = input slot. */
void or_orient(int etype, const double *coords, int *perm);

/* mesher/delaunay.cpp:2408-2535, mesher/delaunay_3d.cpp:2452-2530.
 * coords in local order (nper x dim), C row-major (3x3 or 6x6).  Ke is dense
 * (nper*b) x (nper*b), dof = local_node*b + axis. */
void or_element_matrix(int etype, const double *coords, const double *C, double *Ke);

/* solvers/assembly.cpp:445-537 (2D) / 829-922 (3D): returns nnzb.  With
 * row_size/column_index NULL only counts.  nrows = nnodes. */
int64_t or_pattern(int etype, int64_t nelem, const int32_t *conn, int64_t nnodes,
                   uint32_t *row_size, uint32_t *column_index);

/* solvers/assembly.cpp:657-735 / 1060-1206: serial Kahan scatter of every
 * element matrix (scale 1) into array (zeroed first).  cmat: nphase C
 * matrices; damage (nullable): C_e = C*(1-d_e). */
void or_assemble(int etype, int64_t nelem, const int32_t *conn, const double *coords,
                 const double *cmat, const int32_t *phase, const double *damage,
                 int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
                 double *array);

/* or_assemble with a DOF map: element node i of geometric node conn[e][i]
 * contributes to block row dofmap[conn[e][i]] (periodic condensation; the
 * reference realises it with points that share an id).  dofmap NULL =
 * identity. */
void or_assemble_dofmap(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap, const double *coords,
                        const double *cmat, const int32_t *phase, const double *damage, int64_t nrows,
                        const uint32_t *row_size, const uint32_t *column_index, double *array);
/* or_assemble_dofmap where an element with damage >= fractured_at keeps
   C * residual instead of C * (1 - d)
   (physics/damagemodels/fiberbasedisotropiclineardamage.cpp:37-46). */
void or_assemble_fractured(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap, const double *coords,
                           const double *cmat, const int32_t *phase, const double *damage, double fractured_at,
                           double residual, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
                           double *array);

/* Load of an imposed macro strain E (Voigt, engineering shear): u_E = E.x at
 * every node, f -= sum_e K_e u_E,e scattered through the DOF map. */
void or_macro_rhs(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap, const double *coords,
                  const double *cmat, const int32_t *phase, const double *damage, const double *macro,
                  int64_t nrows, double *f);

/* or_element_strain through a DOF map, plus the macro strain (nullable). */
void or_element_strain_dofmap(int etype, int64_t nelem, const int32_t *conn, const int32_t *dofmap,
                              const double *coords, const double *u, const double *macro, double *strain);

/* Strain (and stress, nullable) at every Gauss point of each element:
   ElementState::getField at the element's Gauss points
   (elements/integrable_entity.cpp:977-1097,1379-1394).  nelem x ngp x 3|6. */
void or_gauss_fields(int etype, int64_t nelem, const int32_t *conn, const double *coords, const double *cmat,
                     const int32_t *phase, const double *damage, const double *u, double *strain, double *stress);

/* solvers/assembly.cpp:125-322: symmetric elimination, unit diagonal,
 * RHS lift, b[id] = g.  dofs need not be sorted. */
void or_apply_dirichlet(int stride, int64_t nrows, const uint32_t *row_size,
                        const uint32_t *column_index, double *array, double *rhs,
                        int64_t nbc, const int64_t *dofs, const double *vals);

/* sparse/sparse_matrix.h:203-333 + sparse/sparse_matrix.cpp:462-547:
 * y = A x - b (b nullable -> y = A x), Kahan per row, OpenMP over rows. */
void or_spmv(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
             const double *array, const double *x, const double *b, double *y);

/* utilities/matrixops.h:246-262: per-thread Kahan + reduction. */
double or_dot(const double *a, const double *b, int64_t n);

/* sparse/sparse_matrix.cpp:216-231 */
void or_inverse_diagonal(int stride, int64_t nrows, const uint32_t *row_size,
                         const uint32_t *column_index, const double *array, double *dinv);
/* solvers/inversediagonal.cpp:84-133; out: (n/2) row-major 2x2 inverses. */
int or_inverse_2x2_diagonal(int stride, int64_t nrows, const uint32_t *row_size,
                            const uint32_t *column_index, const double *array, double *blocks);

/* solvers/conjugategradient.cpp:69-318 with the InverseDiagonal
 * preconditioner.  Returns 1 converged, 0 not.  nit_out = inner iterations. */
int or_cg_reference(int stride, int64_t nrows, const uint32_t *row_size,
                    const uint32_t *column_index, const double *array, const double *b,
                    const double *x0, double eps, int64_t maxit, int nssor,
                    double *x, int64_t *nit_out);

/* Textbook preconditioned CG stopped on ||b - A x|| / ||b|| <= rtol (true
 * residual checked at exit).  precond 0 = point Jacobi, 1 = b x b block
 * Jacobi.  Returns 1 converged.  Used as the FP64 parity solution and as the
 * CPU baseline. */
int or_pcg(int stride, int64_t nrows, const uint32_t *row_size, const uint32_t *column_index,
           const double *array, const double *b, const double *x0, double rtol, int64_t maxit,
           int precond, double *x, int64_t *nit_out, double *relres_out);

/* elements/integrable_entity.cpp:977-1097: element strain (Voigt, engineering
 * shear) from the nodal displacement vector u (dof = node*b + axis), at the
 * element centroid with the reference's centroid Jinv.  out: nelem x 3|6. */
void or_element_strain(int etype, int64_t nelem, const int32_t *conn, const double *coords,
                       const double *u, double *strain);

/* Damage law of the 3D incremental-loading case: Mazars equivalent strain
 * (physics/fracturecriteria/mazars.cpp:124-139 tension branch) of the
 * principal strains, kappa = max(kappa, eps_eq), d = max(d0,
 * 1 - (k0/kappa) exp(-(kappa-k0)/kf)) for kappa > k0, capped at dmax.
 * Updates kappa and d in place.  Returns number of elements whose d grew. */
int64_t or_damage_update(int64_t nelem, int dim, const double *strain, const double *d0,
                         double kappa0, double kappaf, double dmax, double *kappa, double *d);

/* Volume average of C_e (strain) over elements (stress average), with
 * element measure |det J|/2 (tri) or /6 (tet).  out: 3|6 components. */
void or_average_stress(int etype, int64_t nelem, const int32_t *conn, const double *coords,
                       const double *cmat, const int32_t *phase, const double *strain,
                       double *out, double *measure_out);

/* Non-local (smoothed) element strains of linear elements (P1, Tet4): the
 * reference's Mesh::getSmoothedFields, non-space-time branch
 * (mesher/mesh.h:1384-1486), over the neighbourhoods and weights of
 * Mesh::generateCache (mesher/mesh.h:546-640) as FractureCriterion::
 * initialiseCache sets them up (physics/fracturecriteria/fracturecriterion.cpp:
 * 171-190).  For element e (centroid c_e, circumradius r_e):
 *   R_e = max(radius, 1.1 r_e); locus: circle / sphere around c_e of radius
 *   2.5 R_e + 1.5 r_e (2.5 = getSmoothingKernelSize, :28-41), containing the
 *   centroids of the neighbours (geometry_2D.cpp:1472-1484 strict, 3D with
 *   the 1e-12 tolerance of geometry_3D.cpp:1551-1566);
 *   kernel (:43-61) on rrn = |x - c_e|^2 / R_e^2: 0 QUARTIC_COMPACT
 *   ((rrn/2 - 1)^2 where rrn/2 <= 1), 1 GAUSSIAN_NONCOMPACT (exp(-rrn));
 *   cache weight of neighbour i (one Gauss point): mean over its vertices
 *   v of max(w(v) * measure_i, 0); getAverageField returns measure_i times
 *   it (elements/integrable_entity.cpp:2026-2133), so
 *   v_i = measure_i * weight_i;
 *   neighbours of another phase (behaviour source) or fractured
 *   (damage >= threshold, damagemodels/isotropiclineardamage.cpp:57-62)
 *   left out; smoothed = sum v_i eps_i / sum v_i (Kahan-added as
 *   mesher/mesh.h:1453-1462), 0 when sum v_i <= 1e-12.
 * Brute force over all element pairs.  strain, out: nelem x 3|6. */
void or_smoothed_strain(int etype, int64_t nelem, const int32_t *conn, const double *coords, const int32_t *phase,
                        const double *damage, double threshold, double radius, int kernel, const double *strain,
                        double *out);

/* toPrincipal (elements/integrable_entity.cpp:475-560). */
void or_principal(int dim, int doubled, const double *v, double *out);

/* NonLocalMazars::grade, tension branch (physics/fracturecriteria/
   mazars.cpp:73-151), per element from its smoothed strain. */
void or_mazars_score(int dim, int64_t nelem, const int32_t *phase, const int32_t *crit, const double *prm,
                     const double *cmat, const double *sbar, const double *damage, double fractured_at, double *score);

/* One pass of IsotropicLinearDamage::step, Newton branch
   (physics/damagemodels/isotropiclineardamage.cpp:64-92,150-162). */
int64_t or_newton_pass(int etype, int64_t nelem, const int32_t *conn, const double *coords, const int32_t *phase,
                       const int32_t *crit, const double *score, double *damage, double radius, double tol,
                       int32_t *changed);

/* One pass of FiberBasedIsotropicLinearDamage::step over the mesh
   (physics/damagemodels/fiberbasedisotropiclineardamage.cpp:58-92). */
int64_t or_fiber_pass(int etype, int64_t nelem, const int32_t *conn, const double *coords, const int32_t *phase,
                      const int32_t *crit, const double *score, double *damage, double radius, double fraction,
                      double fractured_at, double tol, int64_t first_deferred, int32_t *changed, double *nbr_max);

/* FeatureTree::renumber (features/features.cpp:603-729, one time plane):
 * the elements are walked in order and each node takes the next id the first
 * time one of them names it; nodes no element names keep -1.  Returns the
 * number of ids given. */
int64_t or_renumber(int64_t nnodes, int64_t nelem, int nper, const int32_t *conn, int32_t *ids);

/* DelaunayTree::addSharedNodes(1, 1, ...) (mesher/delaunay.cpp:260-355):
 * triangles in order, sides (v0,v1), (v1,v2), (v2,v0); a side takes the
 * mid-edge node an earlier triangle made on the same edge, else a new one at
 * a*0.5 + b*0.5 (a, b in the side's order), numbered nnodes, nnodes + 1, ...
 * as they are made.  conn6 = [v0, m01, v1, m12, v2, m20] per triangle;
 * mid_xy (2 per new node, room for 3*ntri).  Returns the number made. */
int64_t or_add_midpoints(int64_t nnodes, const double *xy, int64_t ntri, const int32_t *conn3, int32_t *conn6,
                         double *mid_xy);

/* VoxelWriter sampling (utilities/writer/voxel_writer.cpp:88-272): grid
 * corners bl/tr = c -/+ 0.5 * (s * (n-1) / n) per axis of the box (centre c,
 * size s); for every tetrahedron in order, the voxels i of the index range
 * [(int)(n*(min-bl)/(tr-bl)), n*(max-bl)/(tr-bl)) of its bounding box, at
 * point bl + (tr-bl)*i/(n-1), whose point or one of the 8 points offset by
 * +-16e-12 lies in it (barycentric weights of vertices 2, 3, 0 >= -1e-12,
 * their sum <= 1 + 3e-12, geometry_3D.cpp:451-533) take it as owner, later
 * tetrahedra overwriting earlier ones.  owner[k + ny*j + ny*nx*i], -1 where
 * none.  box: cx cy cz sx sy sz. */
void or_voxel_owner(int64_t nnodes, const double *xyz, int64_t ntet, const int32_t *conn, const double *box,
                    int nx, int ny, int nz, int32_t *owner);

#ifdef __cplusplus
}
#endif
#endif
