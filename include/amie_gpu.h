/* amie_gpu.h -- C ABI of the MI355X (gfx950) FE-solve hot path for AMIE.
 *
 * The boundary sits where the reference hands a finished linear system to a
 * LinearSolver and where it integrates / assembles element matrices:
 *
 *   LinearSolver::solve(x0, precond, eps, maxit, verbose)   solvers/solver.h:46-52
 *   Assembly::getMatrix() / getForces()                      solvers/assembly.h:337-346
 *   Assembly::mgprepare() + mgsolve(LinearSolver*, ...)      solvers/assembly.cpp:1225-1228,1955-1969
 *   Assembly::add(element, scale) + make_final()             solvers/assembly.cpp:106-123,425-1223
 *   Assembly::setPointAlong(variable, value, node)           solvers/assembly.cpp:1590-1634
 *   DelaunayTriangle/Tetrahedron::getElementaryMatrix        mesher/delaunay.cpp:2408, delaunay_3d.cpp:2452
 *   FeatureTree::stepElements (strain gather + damage)       features/features.cpp:4875-5100
 *
 * All arguments are plain pointers and sizes; host arrays are caller-owned
 * and copied.  Every call returns an int status (AMIE_OK = 0, 1 = solver did
 * not converge, < 0 = error) instead of exiting the process as the reference
 * does (solvers/conjugategradient.cpp:160-165); amie_gpu_last_error() gives
 * the message of the calling thread's last failure.  Handles are not
 * thread-safe: use one handle per host thread or lock around it.
 *
 * Layout conventions:
 *  - element types: AMIE_P1 (3-node triangle), AMIE_P2 (6-node triangle, nodes
 *    [v0, m01, v1, m12, v2, m20]), AMIE_TET4 (4-node tetrahedron), each in
 *    the local order the reference produces after its orientation rule
 *    (amie_mesh_orient applies it);
 *  - DOF = node*b + axis with b = 2 (2D) or 3 (3D);
 *  - the reference block CSR: row_size[nrows], column_index[nnzb],
 *    array[nnzb*b*colLen], blocks column-major, colLen = b + b%2
 *    (sparse/sparse_matrix.h:129-137, sparse/sparse_vector.cpp:27-58).
 */
#ifndef AMIE_GPU_H
#define AMIE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AMIE_OK = 0,
    AMIE_NOT_CONVERGED = 1,
    AMIE_ERR_ARG = -1,
    AMIE_ERR_HIP = -2,
    AMIE_ERR_NOMEM = -3,
    AMIE_ERR_NAN = -4,
    AMIE_ERR_STATE = -5,
    AMIE_ERR_BUSY = -6
};

enum { AMIE_P1 = 1, AMIE_P2 = 2, AMIE_TET4 = 3 };

/* solve precision.  Every mixed mode keeps the residual r = b - K x, the
 * solution and the CG scalars in FP64 (refinement) and iterates on FP32 or
 * FP64 vectors inside:
 *   AMIE_FP64         FP64 matrix and vectors (the iteration reads a scaled
 *                     FP64 copy S K S, built on the first AMIE_FP64 or
 *                     AMIE_MIXED_K64 solve after an assembly: one more FP64
 *                     matrix of device memory)
 *   AMIE_MIXED        FP32 matrix, vectors and SpMV accumulation
 *   AMIE_MIXED64      FP32 matrix, FP64 vectors
 *   AMIE_MIXED_K64    FP64 matrix, FP32 vectors, FP64 accumulation
 *   AMIE_MIXED_SPLIT  matrix as FP32 + bf16 remainder (6 bytes a value,
 *                     relative error < 2^-32), FP32 vectors, FP64
 *                     accumulation; hands a problem over to the FP64
 *                     matrix when its accuracy floor is reached
 *   AMIE_MIXED_FIX48  scaled matrix as 48-bit fixed point (6 bytes a value,
 *                     absolute error < 2^-47: the scaled SPD matrix has
 *                     |Ks_ij| <= 1), FP32 vectors, FP64 accumulation: the
 *                     FP64 iteration count at the split's traffic
 * Problems whose FP32-vector iteration stagnates continue in AMIE_FP64 when
 * opts.escalate is set. */
enum { AMIE_FP64 = 0, AMIE_MIXED = 1, AMIE_MIXED64 = 2, AMIE_MIXED_K64 = 3, AMIE_MIXED_SPLIT = 4, AMIE_MIXED_FIX48 = 5 };
enum { AMIE_JACOBI = 0, AMIE_BLOCK_JACOBI = 1 };      /* preconditioner */
enum { AMIE_REFINE_REPLACE = 0, AMIE_REFINE_RESTART = 1 };
enum { AMIE_ASSEMBLE_GATHER = 0, AMIE_ASSEMBLE_COLOURED = 1 };  /* amie_gpu_batch_set_assembly */

typedef struct amie_gpu_ctx amie_gpu_ctx;
typedef struct amie_gpu_batch amie_gpu_batch;
typedef struct amie_gpu_server amie_gpu_server;

/* One finite-element problem: mesh, materials, damage, boundary conditions
 * and right-hand sides. */
typedef struct {
    int32_t elem_type;       /* AMIE_P1 | AMIE_P2 | AMIE_TET4 */
    int64_t nnodes;          /* geometric nodes */
    int64_t nelem;
    const double *coords;    /* nnodes*dim */
    const int32_t *conn;     /* nelem*nper geometric node ids, oriented */
    const int32_t *dofmap;   /* nullable: geometric node -> DOF node (periodic
                                condensation); NULL = identity */
    int64_t ndofnodes;       /* DOF nodes (block rows); = nnodes when dofmap is NULL */
    int32_t nphase;
    const double *cmat;      /* nphase C matrices, 3x3 (2D) or 6x6 (3D), row-major */
    const int32_t *phase;    /* nullable: element phase, default 0 */
    const double *damage;    /* nullable: element damage d, C_e = C*(1-d) */
    /* Dirichlet conditions (Assembly::setPointAlong): DOF ids and values;
       a later entry for the same DOF replaces an earlier one. */
    int64_t nbc;
    const int64_t *bc_dofs;
    const double *bc_vals;
    /* right-hand sides: nrhs columns of external forces (nullable, ndof each,
       column-major), plus optional macro strains (nrhs x 3|6, Voigt with
       engineering shear) imposed as u = E.x + periodic fluctuation */
    int32_t nrhs;
    const double *forces;
    const double *macro_strain;
} amie_problem_desc;

typedef struct {
    int32_t precision;       /* AMIE_FP64 | AMIE_MIXED | ... (see above) */
    int32_t precond;         /* AMIE_JACOBI | AMIE_BLOCK_JACOBI */
    int32_t refine;          /* AMIE_REFINE_REPLACE | AMIE_REFINE_RESTART */
    double rtol;             /* stop on ||b - K x|| / ||b|| <= rtol (FP64 residual) */
    double refine_drop;      /* FP32 residual drop that triggers an FP64 refinement (mixed) */
    int64_t ckpt_interval;   /* mixed: FP64 refinement at least every this many iterations (<= 0: 128) */
    int64_t maxit;           /* per right-hand side, <= 0: max(2 ndof, 1000) (a mixed-precision solve of a
                                tiny system needs a few refinement cycles beyond ndof iterations) */
    int32_t use_x0;          /* 1: start from the current device solution */
    int32_t use_graph;       /* 1: replay hipGraphs of iteration chunks (default 0: plain launches) */
    int32_t chunk;           /* iterations per chunk between host status checks */
    int32_t time_kernels;    /* 1: time every SpMV-kernel launch with HIP events */
    int32_t escalate;        /* AMIE_MIXED: continue stagnating problems in AMIE_FP64 (default 1) */
    int32_t max_grid;        /* > 0: cap on workgroups of the set-up kernels (0: 16 per CU) */
    double ref_eps;          /* > 0: the reference's absolute stop instead of rtol:
                                sqrt(r . D^-1 r) <= max(1e-12, ref_eps) with D = diag(K)
                                (solvers/conjugategradient.cpp:101,218,305) */
    int32_t autotune;        /* 1: the first solve of a precision / preconditioner on this batch
                                runs amie_gpu_batch_autotune first (default 0) */
    int64_t stall_iters;     /* FP32-vector modes with escalate: a right-hand side whose FP64
                                residual has not halved in this many iterations stalls its problem,
                                which then continues in AMIE_FP64 (<= 0: 8192) */
    int32_t persistent;      /* > 0: AMIE_FP64 and AMIE_MIXED_FIX48 run each solve pass as one
                                persistent launch of at most this many workgroups (all resident),
                                the iteration phases separated by device-wide barriers instead of
                                kernel boundaries; the same results bit for bit.  For small
                                batches and served groups: a barrier's cost grows with the
                                workgroups (about 2-3 us at 64).  0 (default): one launch per
                                phase. */
} amie_solve_opts;

typedef struct {
    int32_t status;          /* AMIE_OK | AMIE_NOT_CONVERGED | AMIE_ERR_NAN */
    int32_t nproblems_converged;
    int64_t iterations_max;  /* max over problems and right-hand sides */
    int64_t iterations_sum;
    int64_t refinements;     /* FP64 residual evaluations, summed */
    double relres_max;       /* max final ||b - K x|| / ||b|| (FP64) */
    double solve_ms;         /* wall time of the solve on the stream */
    double spmv_kernel_ms;   /* summed HIP-event time of the fused SpMV kernel (time_kernels) */
    int64_t spmv_launches;
    double spmv_bytes;       /* algorithmic bytes per SpMV launch summed over launches */
    int32_t nproblems_escalated; /* problems continued in AMIE_FP64 */
    int32_t nproblems_handover;  /* AMIE_MIXED_SPLIT: problems whose SpMV moved to the FP64
                                    matrix at the split matrix's accuracy floor */
    int64_t iterations_handover; /* iterations run after such a hand-over, summed */
    double update_kernel_ms; /* summed HIP-event time of the CG update kernel (time_kernels) */
    int64_t update_launches;
    double update_bytes;     /* algorithmic bytes of the update launches, summed */
} amie_solve_stats;

/* Damage law of the incremental-loading case (Mazars tension branch with
 * B_t = 1/kappa_f, physics/fracturecriteria/mazars.cpp:113-139). */
typedef struct {
    double kappa0;
    double kappaf;
    double dmax;             /* cap of d (residual stiffness 1 - dmax) */
} amie_damage_params;

/* ------------------------------------------------------------ context */
int amie_gpu_create(amie_gpu_ctx **ctx, int device);
void amie_gpu_destroy(amie_gpu_ctx *ctx);
const char *amie_gpu_last_error(void);
/* Device properties useful to callers (admission control). */
int amie_gpu_device_info(amie_gpu_ctx *ctx, size_t *free_bytes, size_t *total_bytes, int *num_cu);
void amie_gpu_default_opts(amie_solve_opts *opts);

/* Achievable HBM bandwidth: a plain float4 stream copy of `bytes` bytes,
 * repeated `iters` times; returns the mean time of one copy (read + write =
 * 2*bytes of traffic).  Also the reference kernel for scaling hardware
 * byte counters. */
int amie_gpu_stream_copy(amie_gpu_ctx *ctx, size_t bytes, int32_t iters, double *ms_per_copy);

/* ------------------------------------------------------ host mesh helpers
 * Synthetic inputs of the benchmark configurations (no GPU needed). */
/* Reference orientation rule (geometry/geometry_2D.cpp:476-497,
 * geometry/geometry_3D.cpp:254-268) applied in place to conn. */
int amie_mesh_orient(int32_t elem_type, int64_t nelem, int32_t *conn, const double *coords);
/* Structured unit-square / unit-cube meshes; sizes via the *_size call. */
int amie_mesh_structured_size(int32_t elem_type, int32_t n, int64_t *nnodes, int64_t *nelem);
int amie_mesh_structured(int32_t elem_type, int32_t n, double *coords, int32_t *conn);
/* Periodic DOF map of the structured P1 square: node (i,j) -> (i mod n, j mod n). */
int amie_mesh_periodic_dofmap(int32_t n, int32_t *dofmap);
/* placement2D (utilities/placement.cpp:87-162) for circular inclusions in an
 * axis-aligned rectangle (width, height, centre), no pre-placed inclusions
 * or exclusion zones: radii in the order given (the reference's callers sort
 * them largest first), gap min_dist, one budget of tries_max tries for the
 * whole list, rotation range orientation (drawn, ignored by circles), seed
 * of std::default_random_engine.  Same placements as the reference, bit for
 * bit: placed_index (n), placed_xyr (3n: centre, final radius), nplaced. */
int amie_place_inclusions_2d(double width, double height, double cx, double cy, int64_t n, const double *radii,
                             double min_dist, int64_t tries_max, double orientation, uint64_t seed,
                             int64_t *placed_index, double *placed_xyr, int64_t *nplaced);
/* Seeded rejection sampling of circular inclusions in the unit square
 * (fully inside, pairwise gap >= gap) until area fraction >= target or the
 * try cap; returns the number placed in *ncircles (xyr: 3 per circle). */
int amie_inclusions_2d(uint64_t seed, double target_fraction, double rmin, double rmax, double gap,
                       int64_t max_circles, double *xyr, int64_t *ncircles, double *fraction);
/* Phase of every element by centroid: 1 inside any circle, else 0. */
int amie_phase_by_centroid(int32_t elem_type, int64_t nelem, const int32_t *conn, const double *coords,
                           int64_t ncircles, const double *xyr, int32_t *phase);
/* Seeded Gaussian random field on an n^3 voxel grid (correlation length
 * ell, separable Gaussian filter of white noise), mapped affinely to
 * [lo, hi]; one value per tetrahedron of the 6-tet split. */
int amie_damage_field_3d(uint64_t seed, int32_t n, double ell, double lo, double hi, double *d0);
/* Element stiffness of one element on the host, with exactly the
 * formulation the assembly kernel uses (Jinv at the centroid, reference Gauss
 * rules, K_ij = sum_g w_g B_i C B_j^T): DelaunayTriangle /
 * DelaunayTetrahedron::getElementaryMatrix, mesher/delaunay.cpp:2408-2535.
 * coords: nper x dim in local order; Ke: (nper*b)^2, dof = node*b + axis. */
int amie_element_matrix(int32_t elem_type, const double *coords, const double *C, double *Ke);
/* Isotropic C matrix (utilities/tensor.cpp:1845-1931, plane stress in 2D). */
void amie_cauchy_green(double E, double nu, int32_t dim, double *C);

/* Batched element integration on the device (no assembly): element e has
 * coordinates coords[e][nper][dim] and material cmat[e] (3x3 | 6x6);
 * Ke[e] is (nper*b)^2 as amie_element_matrix. */
int amie_gpu_element_matrices(amie_gpu_ctx *ctx, int32_t elem_type, int64_t nelem, const double *coords,
                              const double *cmat, double *Ke);

/* --------------------------------------------------- mesh input (device)
 * Meshes as the reference's FeatureTree hands them over: elements in its
 * mesh iteration order.  Host arrays in and out; device_ms (nullable) is the
 * device time of the work with the inputs already in HBM (HIP events).
 *
 * FeatureTree::renumber (features/features.cpp:603-729): walking the
 * elements in order, each node gets the next id the first time an element
 * names it; ids[nnodes] (-1: named by no element), count = ids given. */
int amie_gpu_renumber(amie_gpu_ctx *ctx, int64_t nnodes, int64_t nelem, int32_t nper, const int32_t *conn,
                      int32_t *ids, int64_t *count, double *device_ms);
/* DelaunayTree::addSharedNodes(1, 1, ...) (mesher/delaunay.cpp:260-355), the
 * mid-edge nodes FeatureTree::stitch adds for QUADRATIC triangles: side s of
 * triangle t (vertices conn3[t][s], conn3[t][(s+1)%3]) takes the mid-edge
 * node an earlier triangle made on the same edge, else a new one at
 * a*0.5 + b*0.5.  New nodes are numbered nnodes, nnodes + 1, ... as they are
 * made; conn6[t] = [v0, m01, v1, m12, v2, m20]; mid_xy needs room for
 * 3*ntri nodes (2 doubles each), nmid = nodes made.  Then amie_gpu_renumber
 * over conn6 gives FeatureTree's final numbering.  Edges are identified by
 * their end nodes (the reference compares positions, which is the same on a
 * valid mesh); the projection of mid-edge nodes onto curved feature
 * boundaries (FeatureTree::projectTrianglesOnBoundaries) is not applied. */
int amie_gpu_add_midpoints(amie_gpu_ctx *ctx, int64_t nnodes, const double *xy, int64_t ntri, const int32_t *conn3,
                           int32_t *conn6, double *mid_xy, int64_t *nmid, double *device_ms);

/* ------------------------------------------------------- result files
 * TriangleWriter text output (utilities/writer/triangle_writer.cpp:31-86,
 * 407-445, 578-608, 1024-1075, 1461-1475), for the reference's viewers:
 * "TRIANGLES", ntri, 3, 2 + total field components, then one line per
 * triangle: vertex coordinates x0 y0 x1 y1 x2 y2, displacements u0x u1x u2x
 * u0y u1y u2y, then for each field, component by component, its values at
 * vertices 0, 1, 2; every value as a C++ stream prints a double, followed by
 * a space.  conn: ntri x nper (vertices at local 0, nper/3, 2nper/3), xy: 2
 * per node for nnodes nodes; the displacements uv and every field are values at the
 * vertices, as the reference evaluates them there: uv ntri x 3 x 2,
 * fields[f] ntri x 3 x sizes[f] (triangle, vertex, component), e.g. the
 * nodal solution at each vertex, and amie_gpu_batch_element_strain's
 * constant strain repeated at the three vertices for TOTAL_STRAIN_FIELD.
 * Host only. */
int amie_write_triangles(const char *path, int64_t ntri, int32_t nper, const int32_t *conn, int64_t nnodes,
                         const double *xy, const double *uv, int32_t nfield, const int32_t *sizes,
                         const double *const *fields);
/* MultiTriangleWriter (triangle_writer.cpp:73-86, 538-573, 1493-1500): the
 * index file `head` starts with "MULTI"; each append writes <base>_<step>
 * (step 1, 2, ... as the caller counts) in the TriangleWriter format and
 * adds that name to the index file.  Host only. */
int amie_multi_triangles_begin(const char *head);
int amie_multi_triangles_append(const char *head, const char *base, int32_t step, int64_t ntri, int32_t nper,
                                const int32_t *conn, int64_t nnodes, const double *xy, const double *uv,
                                int32_t nfield, const int32_t *sizes, const double *const *fields);
/* BinaryTriangleWriter (triangle_writer.cpp:447-478, 1477-1491), same
 * arguments: header "BIN_TRIANGLES", the column count, 3, -1 (the
 * reference's (layers - 6) / 3 for one layer); per triangle the six
 * coordinates, then every later column normalised over the triangles to
 * 1..255 as the voxel files are, written as integers.  Host only. */
int amie_write_triangles_bin(const char *path, int64_t ntri, int32_t nper, const int32_t *conn, int64_t nnodes,
                             const double *xy, const double *uv, int32_t nfield, const int32_t *sizes,
                             const double *const *fields);

/* VoxelWriter sampling on the device (utilities/writer/voxel_writer.cpp:
 * 88-272) for a Tet4 mesh with one value set per element (a Tet4 field is
 * constant over the element, e.g. REAL_STRESS_FIELD): an nx x ny x nz grid
 * spans c -/+ 0.5 s (n-1)/n per axis of the box (box: cx cy cz sx sy sz);
 * walking the tetrahedra in order, each voxel of the index range of a
 * tetrahedron's bounding box whose point (or one of the 8 points offset by
 * +-16e-12) lies in it takes its values, later tetrahedra overwriting
 * earlier ones (on the device: the largest owning index).  Barycentric test
 * of geometry_3D.cpp:451-533 evaluated by volumes.  values: ncomp x
 * (nx*ny*nz), voxel index k + ny*j + ny*nx*i, -1e9 where no tetrahedron
 * owns the voxel; that index numbers cubic grids only, so nx = ny = nz is
 * required.  device_ms (nullable): device time with inputs in HBM. */
int amie_gpu_voxelize(amie_gpu_ctx *ctx, int64_t nnodes, const double *xyz, int64_t ntet, const int32_t *conn,
                      int32_t ncomp, const double *field, const double *box, int32_t nx, int32_t ny, int32_t nz,
                      double *values, double *device_ms);
/* amie_gpu_voxelize on device arrays, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  xyz (nnodes x 3), conn (ntet x 4), field (ntet x ncomp) and
 * values (ncomp x nx*ny*nz) are device pointers the caller owns; box is a
 * host array.  The voxel ownership goes to a static device array of
 * AMIE_VOXEL_SCRATCH_MAX int32 entries (16 MiB) that each GPU the library
 * runs on holds until the process ends; larger grids are refused with
 * AMIE_ERR_ARG.  Work is queued on `stream` (a hipStream_t; NULL: the
 * context's stream) and the call returns without waiting: values are ready
 * when the stream reaches that point.  Every call on one GPU shares its
 * array (each call fills it before use), so calls that overlap on one GPU
 * race.  Tetrahedra naming a node outside [0, nnodes) own no voxel. */
#define AMIE_VOXEL_SCRATCH_MAX (1 << 22)
int amie_gpu_voxelize_noalloc(amie_gpu_ctx *ctx, void *stream, int64_t nnodes, const double *xyz, int64_t ntet,
                              const int32_t *conn, int32_t ncomp, const double *field, const double *box, int32_t nx,
                              int32_t ny, int32_t nz, double *values);
/* VoxelWriter files (voxel_writer.cpp:52-86, 433-442, 548-571): component c
 * of values to <base>_<nx>_<c>.vox, header "VOXELS", 1, nx, ny, nz, then one
 * byte per voxel: the value mapped to [1, 255] over the symmetric range set
 * by the 1 % and 99 % points of the sorted component.  Host only. */
int amie_write_voxels(const char *base, int32_t nx, int32_t ny, int32_t nz, int32_t ncomp, const double *values);

/* ------------------------------------------------------------- batches
 * A batch holds one or more problems resident in HBM: pattern, element->
 * block maps, FP64 and FP32 matrices, damage state and Krylov vectors. */
int amie_gpu_batch_create(amie_gpu_ctx *ctx, int32_t nprob, const amie_problem_desc *probs,
                          amie_gpu_batch **out);
/* Drop-in for a matrix the reference assembled itself (Assembly::getMatrix()
 * and getForces()): one problem, one right-hand side. */
int amie_gpu_batch_from_bcsr(amie_gpu_ctx *ctx, int32_t stride, int64_t nrows, const uint32_t *row_size,
                             const uint32_t *column_index, const double *array, const double *rhs,
                             amie_gpu_batch **out);
/* New matrix values and right-hand side for a batch made by
 * amie_gpu_batch_from_bcsr, in the same pattern (row_size and column_index
 * as then): the values are re-uploaded and re-scaled, the layout and the
 * device buffers are kept.  What a solver reused across
 * Assembly::mgsolve calls needs when only values change (damage steps,
 * new loads: solvers/assembly.cpp:1955-1969). */
int amie_gpu_batch_update_bcsr(amie_gpu_batch *b, const double *array, const double *rhs);
void amie_gpu_batch_destroy(amie_gpu_batch *b);
/* HBM bytes held by the batch. */
size_t amie_gpu_batch_bytes(const amie_gpu_batch *b);
/* Element integration + scatter + Dirichlet elimination + scaling, all on
 * the device (re-run after damage changes). */
int amie_gpu_batch_assemble(amie_gpu_batch *b);
/* Scatter method of the assembly (solvers/assembly.cpp:657-735 is the serial
 * loop both replace), then re-assembles.  AMIE_ASSEMBLE_GATHER (default): one
 * thread per block, its element contributions added in the reference's order
 * with its Kahan compensation; every block written once.  AMIE_ASSEMBLE_COLOURED:
 * one launch per element colour, each element adding its blocks (no two
 * elements of a colour share a block); the sums are ordered by colour.  Both
 * are deterministic; the gather needs fewer than 2^31 element blocks in the
 * batch. */
int amie_gpu_batch_set_assembly(amie_gpu_batch *b, int32_t method);
int amie_gpu_batch_solve(amie_gpu_batch *b, const amie_solve_opts *opts, amie_solve_stats *stats);
/* Outcome of problem p's last solve: status (AMIE_OK, AMIE_NOT_CONVERGED,
 * AMIE_ERR_NAN), iterations, refinements and the largest final relative
 * residual over its right-hand sides. */
int amie_gpu_batch_problem_info(amie_gpu_batch *b, int32_t p, int32_t *status, int64_t *iterations,
                                int64_t *refinements, double *relres);
/* Solution of problem p, original DOF order, nrhs columns (ndof each). */
int amie_gpu_batch_get_solution(amie_gpu_batch *b, int32_t p, double *x);
int amie_gpu_batch_set_solution(amie_gpu_batch *b, int32_t p, const double *x);
/* Pattern and assembled matrix of problem p in the reference layout (after
 * Dirichlet elimination) and its right-hand sides. */
int amie_gpu_batch_get_nnzb(amie_gpu_batch *b, int32_t p, int64_t *nrows, int64_t *nnzb);
int amie_gpu_batch_get_matrix(amie_gpu_batch *b, int32_t p, uint32_t *row_size, uint32_t *column_index,
                              double *array, double *rhs);
/* Element -> block slot map of problem p: nelem*nper*nper indices into the
 * reference block order (column_index positions), pair (i,j) row-major. */
int amie_gpu_batch_get_element_blocks(amie_gpu_batch *b, int32_t p, int64_t *slots);
/* y = K x for problem p (FP64 matrix, or the scaled FP32 matrix un-scaled
 * back when precision = AMIE_MIXED); ndof vectors, original order. */
int amie_gpu_batch_spmv(amie_gpu_batch *b, int32_t p, int32_t precision, const double *x, double *y);
/* Times the iteration's SpMV kernel (q = Ks p with its dot products) over
 * every problem of the batch, `iters` launches after one warm-up: the mean
 * launch time and the algorithmic bytes of one launch (values + the column
 * indices the kernel loads + row pointers + p, r read and q written once;
 * padding not counted).  Uses the autotuned kernel when there is one.  The
 * solver state is left as it was. */
int amie_gpu_batch_time_spmv(amie_gpu_batch *b, int32_t precision, int32_t precond, int32_t iters,
                             double *us_per_launch, double *bytes_per_launch);
/* Autotuner of the iteration's SpMV kernel (no reference counterpart: the
 * reference's SpMV, sparse/sparse_matrix.h:203-678, has one fixed form).
 * Times each candidate, `iters` launches after one warm-up with every
 * problem of the batch active, and keeps the fastest for later solves of
 * this batch with the same precision and preconditioner.  Candidates: half,
 * equal and twice the default number of blocks per step of the row loop,
 * each with four tile orders: XCD-contiguous ranges (0), or chunks of C
 * tiles dealt round-robin over the 8 XCDs (C = 1: launch order, 8, 64);
 * with AMIE_MIXED_FIX48 also the value-indexed matrix (format 1: a 16-bit
 * index per block into a table of the distinct FIX48 blocks, built on the
 * device when there are at most 65536 of them, then rebuilt at every
 * assembly).  Each candidate is timed twice (iters launches each time, in
 * two rounds over the list) and keeps the faster time.  Results do not
 * depend on the choice (the same values, the sums in the same order).
 * Precisions with candidates: AMIE_FP64,
 * AMIE_MIXED and AMIE_MIXED_FIX48.  Outputs (each nullable): the chosen
 * blocks per step, tile order and format (0 value planes, 1 value-indexed)
 * and its time per launch. */
int amie_gpu_batch_autotune(amie_gpu_batch *b, int32_t precision, int32_t precond, int32_t iters,
                            int32_t *steps, int32_t *tile_order, int32_t *format, double *us_per_launch);
/* Sets the SpMV candidate directly (a choice kept from an earlier
 * amie_gpu_batch_autotune, for instance); steps must be one of the
 * candidates', tile_order 0 or a chunk of 1..4096 tiles, format 1 only with
 * AMIE_MIXED_FIX48 and a matrix of at most 65536 distinct blocks
 * (AMIE_ERR_ARG otherwise).  No reference counterpart. */
int amie_gpu_batch_set_spmv_kernel(amie_gpu_batch *b, int32_t precision, int32_t precond, int32_t steps,
                                   int32_t tile_order, int32_t format);
/* Element strains (nelem x 3|6, Voigt, engineering shear) of right-hand side
 * rhs, including the macro strain. */
int amie_gpu_batch_element_strain(amie_gpu_batch *b, int32_t p, int32_t rhs, double *strain);
/* Strain and stress (nullable) at every Gauss point of problem p's elements
 * for right-hand side rhs: nelem x ngp x 3|6 each, ngp = 4 for P2 (the
 * 4-point rule of elements/elements.cpp:684-689), 1 for P1 and Tet4.  Restates
 * ElementState::getField(TOTAL_STRAIN_FIELD / REAL_STRESS_FIELD) at the
 * Gauss points (elements/integrable_entity.cpp:977-1097,1379-1394): the
 * shape-function derivatives at the point with the inverse Jacobian cached
 * at the centroid; stress = strain . C (1 - d). */
int amie_gpu_batch_gauss_fields(amie_gpu_batch *b, int32_t p, int32_t rhs, double *strain, double *stress);
/* Volume-averaged stress per right-hand side: nprob x nrhs x 3|6.  For unit
 * macro strains this is the effective stiffness. */
int amie_gpu_batch_average_stress(amie_gpu_batch *b, double *out);
/* One damage step on the device for every problem: strain of the current
 * solution (right-hand side 0), kappa history, d update, then re-assembly.
 * d0 (nullable, per problem element count) seeds the damage on first call.
 * n_grown (nullable) returns the number of elements whose damage grew. */
int amie_gpu_batch_damage_step(amie_gpu_batch *b, const amie_damage_params *prm, int64_t *n_grown);
/* Non-local damage driving strain (linear elements, P1 and Tet4): restates
 * the reference's smoothed fields, Mesh::getSmoothedFields
 * (mesher/mesh.h:1384-1486) over the neighbourhoods of Mesh::generateCache
 * (mesher/mesh.h:546-640) as FractureCriterion::initialiseCache builds them
 * (physics/fracturecriteria/fracturecriterion.cpp:171-190).  radius is the
 * material characteristic radius (the reference default is 0.008; <= 0
 * turns the smoothing off), kernel 0 QUARTIC_COMPACT (the reference
 * default) or 1 GAUSSIAN_NONCOMPACT, elements with damage >= threshold count
 * as fractured and are left out (the reference default threshold is 1).
 * Once set, amie_gpu_batch_damage_step drives the damage law with the
 * smoothed strains; amie_gpu_batch_event_step runs the reference's
 * event-driven iteration on them.  Each neighbour counts with its measure
 * times its cache weight (ElementState::getAverageField returns both,
 * elements/integrable_entity.cpp:2026-2133). */
int amie_gpu_batch_set_nonlocal(amie_gpu_batch *b, double radius, int32_t kernel, double threshold);
/* The smoothed strains of problem p for right-hand side rhs (nelem x 3|6). */
int amie_gpu_batch_smoothed_strain(amie_gpu_batch *b, int32_t p, int32_t rhs, double *strain);
int amie_gpu_batch_get_damage(amie_gpu_batch *b, int32_t p, double *d, double *kappa);
int amie_gpu_batch_set_damage(amie_gpu_batch *b, int32_t p, const double *d);
/* Event-driven damage: one FeatureTree::step (features/features.cpp:
 * 6150-6373, stepElements :4875-5290) of elements with the behaviour
 * StiffnessAndFracture(C, NonLocalMazars(threshold, E, nu, Gf, radius)) and
 * its default damage model FiberBasedIsotropicLinearDamage (physics/
 * stiffness_and_fracture.cpp:24-51, physics/damagemodels/
 * fiberbasedisotropiclineardamage.cpp:14-100), at the current Dirichlet
 * values: repeat { assemble with the current damage (C (1 - d), C * residual
 * once fractured); solve; smooth the strains (amie_gpu_batch_set_nonlocal's
 * restatement, fractured elements left out); score every element
 * (NonLocalMazars::grade, physics/fracturecriteria/mazars.cpp:73-151); one
 * pass in the reference's visiting order in which an unfractured element
 * scoring above 0 and within score_tolerance of the largest score of its
 * neighbourhood (FractureCriterion::getMaxScoreInNeighbourhood,
 * fracturecriterion.cpp:251-280) takes d += fibre_fraction } until a pass
 * changes nothing.  Linear elements (P1, Tet4), right-hand side 0.  With
 * opts.persistent 0 the solves take shorter-latency paths, in FP64
 * whatever opts.precision asks (same stopping rule: FP64 relative residual
 * <= opts.rtol), for one right-hand side: each problem's whole CG in one
 * workgroup when every problem has at most 3 072 block rows (2 048 for
 * Tet4), or a batch of one problem of at most 131 072 block rows over a
 * cooperative grid of up to 64 workgroups; otherwise a batch of at most 32
 * tiles (8 192 block rows) solves with persistent launches (the launch
 * path's results bit for bit). */
typedef struct {
    int32_t nphase;              /* materials the two arrays below cover (>= every problem's nphase) */
    const int32_t *criterion;    /* per material: 1 = StiffnessAndFracture + NonLocalMazars, 0 = elastic */
    const double *mazars;        /* per material: threshold, E, nu, Gf (NonLocalMazars constructor) */
    double radius;               /* material characteristic radius of the criterion */
    int32_t kernel;              /* smoothing kernel: 0 QUARTIC_COMPACT (default), 1 GAUSSIAN_NONCOMPACT */
    double fibre_fraction;       /* damage per event (0.1, the model's default) */
    double fractured_at;         /* thresholdDamageDensity (0.6) */
    double residual;             /* residualStiffnessFraction (1e-4) */
    double score_tolerance;      /* FractureCriterion::scoreTolerance (2e-3) */
    int64_t first_deferred;      /* the pass visits element positions >= first_deferred, then the others:
                                    65 is the order of the reference's OpenMP task pass on one thread
                                    (libgomp runs the first 65 tasks of the region last) */
    int64_t max_passes;          /* stop after this many solves (FeatureTree's own cap only counts
                                    passes without a change) */
    int32_t damage_model;        /* AMIE_DAMAGE_FIBRE (default) or AMIE_DAMAGE_ISOTROPIC_LINEAR */
    int64_t max_passes_per_step; /* FeatureTree::maxitPerStep (200): with AMIE_DAMAGE_ISOTROPIC_LINEAR a
                                    pass that changes nothing ends the step only when no element scores
                                    above 0 or more than this many passes have run */
} amie_event_params;

/* The damage model of the criterion's materials.  AMIE_DAMAGE_FIBRE:
 * FiberBasedIsotropicLinearDamage as above.  AMIE_DAMAGE_ISOTROPIC_LINEAR:
 * IsotropicLinearDamage's Newton branch (physics/damagemodels/
 * isotropiclineardamage.cpp:64-92, postProcess :150-162): T = the largest
 * score of the pass over the problem's criterion elements; an element with
 * score s > 0 and T - s <= 4 score_tolerance is in the damaging window; m =
 * the largest score of its neighbourhood's elements outside the window
 * (max(0, T) when there are none), delta = max(s - m, T - s); unless |T| <
 * score_tolerance / 2, a window element with s > score_tolerance / 20 takes
 * d += clamp(min(delta, s (1 - d)), 0, 1), capped at 1, after the pass.  The
 * model never fractures: the stiffness is C (1 - d), fibre_fraction,
 * fractured_at, residual and first_deferred are not used.  In a batch the
 * passes run until every problem's step has ended. */
#define AMIE_DAMAGE_FIBRE 0
#define AMIE_DAMAGE_ISOTROPIC_LINEAR 1

typedef struct {
    int64_t passes;              /* solves, each followed by one damage pass */
    int64_t cg_iterations;       /* CG iterations over the passes (largest problem per pass) */
    int64_t events;              /* damage increments */
    int64_t fractured;           /* elements with damage >= fractured_at after the step */
    int32_t status;              /* AMIE_OK, or the first failing solve's status */
    int32_t converged;           /* 1: the last pass changed nothing; 0: max_passes reached */
    double solve_ms, damage_ms;  /* wall time in the solves and in the rest of the loop */
} amie_event_stats;

void amie_gpu_default_event_params(amie_event_params *p);
int amie_gpu_batch_event_step(amie_gpu_batch *b, const amie_event_params *prm, const amie_solve_opts *opts,
                              amie_event_stats *st);
/* The fracture scores of problem p's elements after the last pass. */
int amie_gpu_batch_get_scores(amie_gpu_batch *b, int32_t p, double *score);
/* Scale the Dirichlet values of every problem (incremental loading). */
int amie_gpu_batch_scale_bc(amie_gpu_batch *b, double factor);
/* New values for Dirichlet conditions of problem p, each DOF one that the
 * problem was created with (Assembly::setPointAlong on an existing
 * condition, solvers/assembly.cpp:1590-1634: the new value replaces the
 * old); they take effect at the next assembly. */
int amie_gpu_batch_set_bc_values(amie_gpu_batch *b, int32_t p, int64_t nbc, const int64_t *bc_dofs,
                                 const double *bc_vals);

/* -------------------------------------------------------------- serving
 * Continuous batching of solve requests: each request is admitted when its
 * HBM footprint fits the budget and a slot is free; admitted requests share
 * every kernel launch of the solve loop, and new ones join between
 * iteration chunks while others are in flight. */
/* status: AMIE_OK, AMIE_NOT_CONVERGED or AMIE_ERR_NAN for a solve that ran;
 * a request that did not run: AMIE_ERR_ARG (a persistent-key request whose
 * mesh or Dirichlet DOFs differ from the resident problem's), AMIE_ERR_NOMEM
 * (its footprint does not fit next to the idle resident problems), or the
 * error of its batch build (amie_gpu_last_error tells which). */
typedef struct {
    uint64_t ticket;
    int32_t status;
    int32_t iterations;
    double relres;
    double submit_ms;        /* host clock at submission */
    double start_ms;         /* admission */
    double done_ms;          /* completion */
} amie_request_result;

int amie_gpu_server_create(amie_gpu_ctx *ctx, size_t hbm_budget, int32_t max_slots, const amie_solve_opts *opts,
                           amie_gpu_server **out);
void amie_gpu_server_destroy(amie_gpu_server *s);
/* Queue a request (copied).  persistent_key != 0 keeps the problem's
 * matrix, damage state and solution resident between requests carrying
 * that key; each such request is one load step in the order of
 * FeatureTree::stepInternal (features/features.cpp:6150-6373): the
 * request's Dirichlet values replace the resident ones (same mesh and
 * Dirichlet DOFs required), the matrix is re-assembled with the damage so
 * far, the solve starts from the previous solution, then the damage law
 * updates every element (physics/fracturecriteria/mazars.cpp:113-139, here
 * with kappa0 1e-4, kappaf 1e-3, dmax 1 - 1e-6; amie_gpu_server_submit_step
 * takes other parameters).  The first request on a key builds the problem
 * from its description (its damage field as the initial damage).  A request
 * that cannot run completes with a negative status (see
 * amie_request_result). */
int amie_gpu_server_submit(amie_gpu_server *s, const amie_problem_desc *desc, uint64_t persistent_key,
                           double submit_ms, uint64_t *ticket);
/* A load step on a persistent key with its own damage law parameters. */
int amie_gpu_server_submit_step(amie_gpu_server *s, const amie_problem_desc *desc, uint64_t persistent_key,
                                const amie_damage_params *prm, double submit_ms, uint64_t *ticket);
/* HBM the server sets aside for a request at admission (an estimate from
 * the mesh size; the built batch's own size replaces it once built). */
int amie_gpu_server_request_bytes(const amie_problem_desc *desc, int64_t *bytes);
/* Frees the resident problem of a key (its HBM leaves the budget's use);
 * AMIE_ERR_BUSY while a request on it is queued or in flight. */
int amie_gpu_server_release(amie_gpu_server *s, uint64_t persistent_key);
/* Damage and kappa history of a key's resident problem (nelem each), as
 * left by its last load step; AMIE_ERR_BUSY while it is in flight. */
int amie_gpu_server_get_damage(amie_gpu_server *s, uint64_t persistent_key, double *d, double *kappa,
                               int64_t nelem);
/* Advance the solve loop by one chunk (admits, iterates, retires). */
int amie_gpu_server_step(amie_gpu_server *s, double now_ms, int32_t *n_active, int32_t *n_queued);
/* Completed requests since the last poll. */
int amie_gpu_server_poll(amie_gpu_server *s, amie_request_result *out, int32_t max, int32_t *n);
/* Solution of a completed request (until the next poll). */
int amie_gpu_server_get_solution(amie_gpu_server *s, uint64_t ticket, double *x, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
