/*
 * femd_hip.h -- the C ABI of the MI355X implementation of FEMd's 2D hot path:
 * P_k element assembly of bilinear and linear forms into CSR, the CSR SpMV, and CG /
 * GMRES / LGMRES with the Jacobi, SSOR and ILU(0) preconditioners, on one GPU (gfx950).
 *
 * The reference has no C ABI: its core is header-only C++ bound to Python by nanobind
 * (python/src/femd_bindings.cpp, femd_bindings_2d.cpp).  Each entry point below names the
 * reference interface it replaces; the Python package femd_amd calls them from the places
 * the reference's Python layer calls its bindings.
 *
 * Conventions
 *   - Every function returns an int status (FEMD_OK = 0); on failure femd_last_error()
 *     gives the message of the calling thread.  No C++ exception crosses this boundary.
 *   - Host pointers are caller-owned and copied in or out before the call returns.
 *     Functions with the suffix _dev take DEVICE pointers (from femd_dev_alloc) and a
 *     stream (NULL: the library's stream); they return once the work is enqueued, except
 *     where a result must be read back (solver infos), which synchronizes the stream.
 *   - CSR layout is FEMd's: 0-based int32 indptr (n+1) and indices (nnz, sorted and unique
 *     within each row), float64 values (sparse/csr.hpp:30-35, 178-185).
 *   - Handles own their device memory.  Calls on different handles may run concurrently
 *     from different threads; calls on one handle must not overlap.
 */
#ifndef FEMD_HIP_H
#define FEMD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
enum {
    FEMD_OK = 0,
    FEMD_EINVAL = 1,      /* bad argument: std::invalid_argument in the reference (ValueError) */
    FEMD_EPATTERN = 2,    /* an element entry outside the pattern: assembly_2d.hpp:157,162,303 */
    FEMD_EZEROPIVOT = 3,  /* zero diagonal or ILU(0) pivot: precond.hpp:68-78, 181 */
    FEMD_ENOTSPD = 4,     /* reserved: CG reports breakdown in its info instead (cg.hpp:213-218) */
    FEMD_EDEVICE = 5,     /* a HIP runtime error, or the HIP device is missing */
    FEMD_ENOMEM = 6       /* device allocation failed */
};

enum { FEMD_PC_NONE = 0, FEMD_PC_JACOBI = 1, FEMD_PC_ILU0 = 2, FEMD_PC_SSOR = 3 };
enum { FEMD_ORDER_NATURAL = 0, FEMD_ORDER_RCM = 1 };
enum { FEMD_DIAG_RIGHT = 0, FEMD_DIAG_LEFT = 1, FEMD_DIAG_ALTERNATE = 2, FEMD_DIAG_CROSSED = 3 };
enum { FEMD_FP64 = 0, FEMD_FP32 = 1 };

typedef struct femd_mesh_s *femd_mesh;
typedef struct femd_space_s *femd_space;
typedef struct femd_csr_s *femd_csr;
typedef struct femd_precond_s *femd_precond;
typedef struct femd_facets_s *femd_facets;
typedef void *femd_stream;          /* a hipStream_t */

/* The thread's last error message (empty after success). */
const char *femd_last_error(void);

/* Select the HIP device and create the library's stream.  Called implicitly by the first
 * entry point that touches the GPU; returns FEMD_EDEVICE when no GPU is visible. */
int femd_init(int device);

/* Device memory for the _dev entry points (the library's caching pool).  The copies run on the
 * library's stream, after the work already queued there, and return when done. */
int femd_dev_alloc(size_t bytes, void **out);
int femd_dev_free(void *p);
int femd_memcpy_htod(void *dst, const void *src, size_t bytes);
int femd_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int femd_stream_sync(femd_stream s);
/* dst = src for n doubles on the device, by a kernel (no counterpart in the reference; the copy
 * whose HBM counters calibrate the SpMV traffic measurement). */
int femd_vec_copy_dev(int64_t n, const double *src, double *dst, femd_stream s);

/* ---------------------------------------------------------------- meshes
 * femd_mesh_rectangle replaces rectangle_mesh (mesh/structured.hpp:338-351, bound at
 * python/src/femd_bindings.cpp:463-472): grid vertex (i, j) is point j(nx+1)+i, cells
 * split by `diag`, boundary segments with markers[4] = bottom, right, top, left.
 * femd_mesh_from_arrays replaces mesh_from_arrays (mesh/structured.hpp:99-221): clockwise
 * triangles are re-oriented, neighbours linked, user segments kept first, and the other
 * boundary edges added with `marker`.  smarks may be NULL (every segment gets `marker`). */
int femd_mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, int diag,
                        const int markers[4], femd_mesh *out);
int femd_mesh_from_arrays(int np, const double *points, int nt, const int *triangles, int ns,
                          const int *segments, const int *smarks, int marker, femd_mesh *out);
/* Quadrilateral meshes (for Q_k spaces).  femd_quad_mesh_from_arrays replaces the QuadMesh
 * constructor (mesh/quad_mesh.hpp:50-103, bound through python/femd/quadmesh.py:130-143):
 * clockwise quads re-oriented, every quad strictly convex, neighbour i across the edge from
 * vertex i to vertex i+1, the listed segments first, then every other boundary edge in quad
 * order with `marker`.  femd_quad_mesh_rectangle replaces rectangle_quad_mesh
 * (python/femd/quadmesh.py:161-170): rectangle_mesh's grid and sides, quad (i, j) = j nx + i
 * with vertices (i, j), (i+1, j), (i+1, j+1), (i, j+1). */
int femd_quad_mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, const int markers[4],
                             femd_mesh *out);
int femd_quad_mesh_from_arrays(int np, const double *points, int nq, const int *quads, int ns, const int *segments,
                               const int *smarks, int marker, femd_mesh *out);
/* 3 for a triangle mesh, 4 for a quadrilateral one */
int femd_mesh_cell_vertices(femd_mesh m, int *nv);
/* out[3]: npoints, ncells, nsegments */
int femd_mesh_sizes(femd_mesh m, int out[3]);
/* points (np x 2), cells and neighbours (ncells x nv), segments (ns x 2), markers (ns) */
int femd_mesh_arrays(femd_mesh m, double *points, int *triangles, int *neighbours, int *segments, int *smarks);
void femd_mesh_free(femd_mesh m);

/* ---------------------------------------------------------------- spaces
 * Replaces LagrangeSpace2D(mesh, k, dirichlet, all_exterior, equispaced)
 * (fe/lagrange_space_2d.hpp:32-67, bound at python/src/femd_bindings_2d.cpp:571-573) on a
 * triangle mesh (k = 1..5), and LagrangeSpaceQ(mesh, k, ...) (fe/lagrange_space_quad.hpp:112-210)
 * on a quadrilateral one (k = 1..4): the numbering of fe/space_2d.hpp:390-526, bit for bit.
 * Only equispaced nodes are built (equispaced = 0 returns FEMD_EINVAL for k >= 3; for k <= 2
 * the two node sets coincide). */
int femd_space_create(femd_mesh m, int degree, int ndirichlet, const int *dirichlet, int all_exterior,
                      int equispaced, femd_space *out);
/* out[6]: ncells, nloc, raw_dim, dim, nedges, npoints */
int femd_space_sizes(femd_space V, int out[6]);
/* cell_dof and cell_raw: ncells x nloc (-1 = eliminated); ad2raw: dim; raw_xy: raw_dim x 2 */
int femd_space_numbering(femd_space V, int *cell_dof, int *cell_raw, int *ad2raw, double *raw_xy);
void femd_space_free(femd_space V);

/* The CSR pattern of the space with itself (cell_pattern, forms/assembly_2d.hpp:36-41;
 * SparsityPattern::from_cells, sparse/csr.hpp:84-125).  Call with indices == NULL to get
 * nnz, then again with indptr (dim+1) and indices (nnz). */
int femd_pattern(femd_space V, int64_t *nnz, int *indptr, int *indices);

/* ---------------------------------------------------------------- coefficients
 * The coefficient of one term, at every quadrature point: either a register program in the
 * instruction set of forms/integrand.hpp:33-46 (4 ints per instruction: op, out, a, b; ops
 * CONST X Y INPUT ADD MUL POW SIN COS EXP LOG TANH SQRT SINH COSH SECH ABS SIGN; normals on ds
 * terms only; INPUT a reads the known field inputs[a]), evaluated on the GPU at the quadrature
 * points, or, when `values` is not NULL, ncells x nq host values laid out [cell * nq + q]
 * (Cache2D order, fe/cache_2d.hpp:98-118). */
typedef struct femd_field_s femd_field;
typedef struct {
    int ninstr;
    const int *code;
    int nreg;
    int nconst;
    const double *consts;
    const double *values;
    int ninputs;                 /* known fields the program reads (INPUT a reads inputs[a]) */
    const femd_field *inputs;
} femd_coefficient;
/* A known field read by a program's INPUT instruction (forms/integrand.hpp:39; the reference's
 * FormAssembler2D::evaluate_fields, forms/assembler_2d.hpp:336-366): a Lagrange function given by
 * its coefficients (dim of its space, host) on a space over the same cells as the form's, read at
 * the quadrature points as its value (code 0) or its x / y derivative (code 1 / 2), summed in the
 * order of Cache2D::eval_multi (fe/cache_2d.hpp:245-288). */
struct femd_field_s {
    femd_space space;
    const double *coefficients;
    int code;
};

/* The program with its registers renamed by liveness, as the GPU runs it (host only; no
 * counterpart in the reference, whose Integrand keeps one register per node): code_out has
 * 4 * ninstr ints; nreg_out receives the registers used. */
int femd_program_compact(const femd_coefficient *c, int *code_out, int *nreg_out);

/* Quadrature of a degree: nq points (the rule of quadrature/triangle.hpp:22-57). */
int femd_quadrature(int degree, int *nq, double *xi, double *eta, double *w);
/* The quadrilateral rule of a degree (fe/cache_2d.hpp:75-84): tensor Gauss-Legendre with
 * max(1, (degree+3)/2) points per direction on [0,1]^2, xi fastest. */
int femd_quadrature_square(int degree, int *nq, double *xi, double *eta, double *w);

/* ---------------------------------------------------------------- CSR matrices
 * femd_csr_create_space: zero values on the space's own pattern (the K of
 * forms.py:2089).  femd_csr_create: from host arrays (SparseMatrix.from_scipy,
 * python/femd/sparse.py:60-70).  The pattern arrays are validated (sorted, unique, in range). */
int femd_csr_create_space(femd_space V, femd_csr *out);
int femd_csr_create(int nrows, int ncols, const int *indptr, const int *indices, const double *data, femd_csr *out);
/* out: nrows, ncols; nnz */
int femd_csr_info(femd_csr A, int out[2], int64_t *nnz);
int femd_csr_download(femd_csr A, int *indptr, int *indices, double *data);
int femd_csr_upload_values(femd_csr A, const double *data);
/* Device pointers of the arrays (valid while A lives). */
int femd_csr_device_arrays(femd_csr A, const int **indptr, const int **indices, double **data);
/* max |a_ij - a_ji| / max |a_ij| (asymmetry, sparse/csr.hpp:631-644), on the GPU. */
int femd_csr_asymmetry(femd_csr A, double *out);
/* The diagonal (CSRMatrix::diagonal, csr.hpp:245-252). */
int femd_csr_diagonal(femd_csr A, double *out);
void femd_csr_free(femd_csr A);

/* ---------------------------------------------------------------- assembly
 * K (zeroed here) = sum_t integral c_t (D^{a_t} phi_i)(D^{b_t} phi_j) dx over the cells,
 * codes 0 value, 1 d/dx, 2 d/dy: FormAssembler2D::assemble_matrix
 * (forms/assembler_2d.hpp:249-266) -> assemble_matrix_2d (forms/assembly_2d.hpp:219-305).
 * K must come from femd_csr_create_space(V).  Every entry is the reference's sum in ascending
 * element order, so the values are the reference's bit for bit on the same coefficients. */
int femd_assemble_matrix(femd_space V, int quad_degree, int nterms, const int *a, const int *b,
                         const femd_coefficient *coeffs, femd_csr K);
/* f (dim, host) = sum_t integral c_t D^{a_t} phi_i dx: FormAssembler2D::assemble_vector
 * (assembler_2d.hpp:166-188) -> assemble_vector_2d (assembly_2d.hpp:173-211). */
int femd_assemble_vector(femd_space V, int quad_degree, int nterms, const int *a,
                         const femd_coefficient *coeffs, double *f);
/* The same into a device vector (dim doubles), on stream s. */
int femd_assemble_vector_dev(femd_space V, int quad_degree, int nterms, const int *a,
                             const femd_coefficient *coeffs, double *f_dev, femd_stream s);
/* The coefficient program at the quadrature points of the space (ncells x nq, host). */
int femd_eval_coefficient(femd_space V, int quad_degree, const femd_coefficient *c, double *out);
/* femd_assemble_matrix / femd_assemble_vector for a dx group that follows another group of the
 * same form: every stored entry continues its sum (K[p] += each element entry in order), the
 * vector adds the group's own sum (f[i] += block[i]) -- how FormAssembler2D combines its groups
 * (forms/assembler_2d.hpp:166-188, 249-266).  No separate entry point in the reference. */
int femd_assemble_matrix_add(femd_space V, int quad_degree, int nterms, const int *a, const int *b,
                             const femd_coefficient *coeffs, femd_csr K);
int femd_assemble_vector_add(femd_space V, int quad_degree, int nterms, const int *a,
                             const femd_coefficient *coeffs, double *f);

/* ---------------------------------------------------------------- boundary facets (ds)
 * The exterior facets of V whose side marker is one of markers[0..nmarkers) (every exterior facet
 * with all != 0), in the space's facet order, with Gauss-Legendre points exact to `degree` along
 * each edge: the point set of Cache2D(V, degree, markers, all) (fe/cache_2d.hpp:124-183), which
 * LagrangeSpace.boundary_cache builds (python/femd/spaces2d.py:265-270) for ds(markers) terms
 * (python/femd/forms.py:2502-2524).  The set keeps V's address: free it before V. */
int femd_facets_create(femd_space V, int degree, int nmarkers, const int *markers, int all, femd_facets *out);
/* out[0] facets, out[1] points per facet. */
int femd_facets_sizes(femd_facets F, int out[2]);
/* Per (facet, point), facet-major: physical point, weight (reference weight x half the edge
 * length), outward unit normal.  Any pointer may be NULL. */
int femd_facets_points(femd_facets F, double *x, double *y, double *w, double *nx, double *ny);
void femd_facets_free(femd_facets F);
/* K += sum_t integral over the facets of c_t (D^{a_t} phi_i)(D^{b_t} phi_j) ds, each facet's
 * element matrix formed and added in ascending facet order (assemble_matrix_2d on a facet cache,
 * forms/assembler_2d.hpp:249-266).  mode 0 zeroes K first (the form's first group), 1 continues
 * K's sums (a later group).  Coefficient programs may read the outward normal (NX, NY). */
int femd_assemble_matrix_facets(femd_facets F, int nterms, const int *a, const int *b,
                                const femd_coefficient *coeffs, femd_csr K, int mode);
/* f (dim, host) += the group's vector sum_t integral c_t D^{a_t} phi_i ds (assemble_vector_2d on a
 * facet cache, then added as forms/assembler_2d.hpp:166-188 does); mode 0 zeroes f first. */
int femd_assemble_vector_facets(femd_facets F, int nterms, const int *a, const femd_coefficient *coeffs,
                                double *f, int mode);
/* The coefficient program at the facet points (nfacets x nq, host), normals available. */
int femd_eval_coefficient_facets(femd_facets F, const femd_coefficient *c, double *out);

/* ---------------------------------------------------------------- SpMV
 * y = A x (CSRMatrix::apply, sparse/csr.hpp:203-215, bound as CSRMatrix.matvec): each row
 * summed left to right in stored order, so the result is the reference's bit for bit.
 * precision FEMD_FP32 uses the float shadow of the values (made on first use). */
int femd_spmv(femd_csr A, const double *x, double *y);
int femd_spmv_dev(femd_csr A, const void *x_dev, void *y_dev, int precision, femd_stream s);
/* The SpMV launch shape the autotuner picked for A in a precision (every shape sums rows in the
 * reference's order), tuning on first use; name receives its name (name_len bytes), ms its
 * measured time.  No counterpart in the reference. */
int femd_spmv_tuning(femd_csr A, int precision, char *name, int name_len, double *ms);
/* The tuning space: femd_spmv_shape_count() shapes, named by femd_spmv_shape_name.  No
 * counterpart in the reference. */
int femd_spmv_shape_count(void);
int femd_spmv_shape_name(int index, char *name, int name_len);
/* Pin A's SpMV in a precision to the named shape (tests, benchmarks); NULL or "" returns it to
 * the autotuner.  FEMD_EINVAL for an unknown name or one not available in that precision.  A
 * pattern's shape is shared by the matrices on it.  No counterpart in the reference. */
int femd_spmv_select(femd_csr A, int precision, const char *name);
/* The dot product of util/omp.hpp:176-192 (fd.ddot), with its summation order. */
int femd_ddot(int64_t n, const double *x, const double *y, double *out);
int femd_ddot_dev(int64_t n, const double *x_dev, const double *y_dev, double *out, femd_stream s);

/* ---------------------------------------------------------------- preconditioners
 * Replaces JacobiPreconditioner, SSORPreconditioner(omega, ordering) and
 * ILU0Preconditioner(ordering) (sparse/precond.hpp:82-205, bound at
 * python/src/femd_bindings_2d.cpp:170-189): built from A (which it copies), applied as
 * z = M^{-1} r.  ordering FEMD_ORDER_RCM builds on P A P^T with the reverse Cuthill-McKee
 * order of sparse/amd.hpp:166-234.  min_pivot (may be NULL) receives ILU(0)'s smallest
 * pivot.  A zero diagonal or pivot returns FEMD_EZEROPIVOT. */
int femd_precond_create(femd_csr A, int kind, int ordering, double omega, femd_precond *out, double *min_pivot);
int femd_precond_apply(femd_precond M, const double *r, double *z);
int femd_precond_apply_dev(femd_precond M, const double *r_dev, double *z_dev, femd_stream s);
/* kind, ordering, n */
int femd_precond_info(femd_precond M, int out[3]);
/* How an ILU(0) / SSOR preconditioner runs its triangular sweeps (no counterpart in the reference):
 * out[0] 1 when each sweep is one workgroup's planned program (else the level-set kernels), out[1..2]
 * the forward / backward steps, out[3..4] their dependency levels, out[5..6] the bytes of their
 * streams, out[7] the stored entries the two sweeps read (L, U and the diagonal). */
int femd_precond_stats(femd_precond M, int64_t out[8]);
/* The triangular sweeps of an ILU(0) (values = its factor L\U) or SSOR(omega) (values = the matrix)
 * preconditioner, z = M^{-1} r on a pattern in the order it is applied, run on the HOST exactly as the
 * GPU runs them: the one-workgroup schedule femd_precond_create plans (dependency levels cut into
 * steps, LDS slots for the z values, the streamed ring) emulated step by step.  stats (may be NULL)
 * receives the forward and backward steps, dependency levels and LDS slots.  FEMD_EINVAL when the
 * schedule does not fit the LDS.  No counterpart in the reference: a check of the planner without a
 * GPU. */
int femd_sweep_host_apply(int n, const int *indptr, const int *indices, const double *values, int kind, double omega,
                          const double *r, double *z, int stats[6]);
/* The RCM order of a pattern: perm[k] = original index of row k (ordering::rcm_order). */
int femd_rcm(int n, const int *indptr, const int *indices, int *perm);
void femd_precond_free(femd_precond M);

/* ---------------------------------------------------------------- Krylov solvers */
typedef struct {
    int converged;
    int iterations;
    int breakdown;              /* CG: p^T A p <= 0 or r^T z <= 0 (cg.hpp:213-218) */
    double residual;            /* true ||b - A x|| / ||b||, recomputed at the end */
    double precond_residual;    /* GMRES: ||M^{-1}(b - A x)|| / ||M^{-1} b|| */
} femd_solve_info;

/* CG (fd.cg, python/femd/krylov.py:178-202 -> cg_solve, femd_bindings.cpp:799-833): M NULL
 * or Jacobi runs the fused cg_jacobi (krylov/cg.hpp:295-350), any other M the general loop
 * (cg.hpp:245-287).  x holds the initial guess and receives the solution.  The iterates are
 * the reference's bit for bit. */
int femd_cg(femd_csr A, femd_precond M, const double *b, double *x, int max_iter, double tol, femd_solve_info *info);
int femd_cg_dev(femd_csr A, femd_precond M, const double *b_dev, double *x_dev, int max_iter, double tol,
                femd_solve_info *info, femd_stream s);

/* femd_cg_dev with no allocation of any kind (no counterpart in the reference: for callers that
 * cannot allocate device or pinned memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and the same iterates, bit for bit, on a
 * work space held in a static device array of 32 MiB per GPU (kept until the process ends), so:
 *   - the matrix has at most 1 048 576 rows (FEMD_EINVAL beyond);
 *   - A's FP64 SpMV launch shape must exist: call femd_spmv_tuning(A, FEMD_FP64, ...), or any
 *     FP64 product or solve on A, once beforehand (FEMD_EINVAL otherwise);
 *   - M is NULL or any preconditioner (its buffers are made by femd_precond_create);
 *   - every call on one GPU uses that GPU's one work space: calls that overlap on a GPU (other
 *     threads, other streams) race and must be serialized by the caller. */
int femd_cg_dev_noalloc(femd_csr A, femd_precond M, const double *b_dev, double *x_dev, int max_iter, double tol,
                        femd_solve_info *info, femd_stream s);

/* GMRES(restart) / LGMRES(restart, k_aug), left preconditioned (fd.gmres / fd.lgmres,
 * krylov.py:148-175 -> krylov_solve, femd_bindings.cpp:784-798 -> krylov::solve,
 * krylov/adapters.hpp:116-147 -> csnewton gmres / lgmres, vendor/csnewton/csnewton.hpp:174-474). */
int femd_krylov(femd_csr A, femd_precond M, const double *b, double *x, int lgmres, int restart, int max_iter,
                double tol, int k_aug, femd_solve_info *info);

/* ---------------------------------------------------------------- mixed precision
 * FP64 iterative refinement around an FP32 CG (the precision pattern; no reference
 * counterpart): r = b - A x in FP64, then A d = r solved in FP32 by Jacobi-PCG to rtol
 * inner_tol, x += d in FP64, until ||b - A x|| / ||b|| <= tol.  info.iterations counts inner
 * CG iterations, info.breakdown the refinement steps. */
int femd_solve_refined(femd_csr A, const double *b, double *x, double tol, double inner_tol, int max_outer,
                       femd_solve_info *info);
int femd_solve_refined_dev(femd_csr A, const double *b_dev, double *x_dev, double tol, double inner_tol,
                           int max_outer, femd_solve_info *info, femd_stream s);

/* ---------------------------------------------------------------- batches
 * Many independent square systems solved together (the batching pattern; no reference
 * counterpart: the reference solves one system per fd.cg call).  The systems are concatenated
 * into one block-diagonal CSR; each keeps its own CG scalars and convergence flag on the device,
 * so uneven sizes share one launch sequence, replayed from a captured hipGraph when use_graph.
 * femd_batch_create copies the members' patterns and values (the members must stay alive for
 * femd_batch_update_values, which re-copies their values after a re-assembly).
 * Vectors are laid out system after system (row_offsets, np + 1 entries).  Every system runs
 * Jacobi-preconditioned CG (on S A S, S = diag^(-1/2)) in FP32 with reliable FP64 updates: when
 * its r.r has dropped by inner_tol^2 its correction goes into x and its FP64 residual replaces the
 * FP32 one; it stops when ||b_p - A_p x_p|| / ||b_p|| <= tol.  iterations[p] (inner iterations)
 * and residuals[p] may be NULL; info.iterations sums the inner iterations, info.breakdown counts
 * the rounds of reliable updates, info.residual is the worst system's.  Each system's result does
 * not depend on the other systems in the batch (fixed tiling, per-system reductions). */
typedef struct femd_batch_s *femd_batch;
int femd_batch_create(int nsystems, const femd_csr *matrices, femd_batch *out);
int femd_batch_info(femd_batch B, int64_t *nrows, int64_t *nnz, int64_t *row_offsets);
int femd_batch_update_values(femd_batch B);
int femd_batch_solve_refined(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer,
                             int use_graph, int *iterations, double *residuals, femd_solve_info *info);
/* The same with the inner iterations in FEMD_FP32 (the _refined entry points) or FEMD_FP64 (S A S,
 * the vectors and every reduction in FP64: the batched counterpart of the plain FP64 path, for
 * systems whose convergence FP32 arithmetic would slow down).  max_outer caps the reliable
 * updates per system. */
int femd_batch_solve(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer, int precision,
                     int use_graph, int *iterations, double *residuals, femd_solve_info *info);
int femd_batch_solve_dev(femd_batch B, const double *b_dev, double *x_dev, double tol, double inner_tol, int max_outer,
                         int precision, int use_graph, int *iterations, double *residuals, femd_solve_info *info,
                         femd_stream s);
int femd_batch_solve_refined_dev(femd_batch B, const double *b_dev, double *x_dev, double tol, double inner_tol,
                                 int max_outer, int use_graph, int *iterations, double *residuals,
                                 femd_solve_info *info, femd_stream s);
/* The value dictionary the last solve in `precision` used: *distinct its entries over all tiles
 * (-1: none, the values did not allow one or it is turned off), *max_per_tile the most distinct
 * values in one tile (-1: not examined). */
int femd_batch_dictionary(femd_batch B, int precision, int64_t *distinct, int *max_per_tile);
void femd_batch_free(femd_batch B);

/* ---------------------------------------------------------------- serving
 * A request loop for independent square SPD systems of uneven sizes (the serving pattern; no
 * reference counterpart).  femd_server_submit queues a request: an assembled matrix A, a
 * right-hand side b_dev and an output x_dev holding the initial guess, all on the device and kept
 * alive by the caller until the request is reported done; *bytes (may be NULL) gets the device
 * memory it will hold while admitted.  Each femd_server_step queues one round of iterations of
 * every running group, admits in arrival order the waiting requests that fit hbm_budget and
 * max_systems as a new group (its own batch, engine and stream, so groups overlap on the device),
 * completes the round -- reliable FP64 updates of paused systems, and every system that has met
 * tol retires at once, its solution copied to its x_dev -- and frees finished groups.  It returns
 * up to max_done finished requests (tag, inner iterations, relative residual, converged); the
 * caller keeps stepping while requests are queued or running.  Results are bit-identical whatever
 * else is admitted alongside.  Inner iterations in `precision` (FEMD_FP32 with FP64 reliable
 * updates, or FEMD_FP64), inner_tol and max_outer as femd_batch_solve. */
typedef struct femd_server_s *femd_server;
int femd_server_create(int64_t hbm_budget, int max_systems, int precision, double tol, double inner_tol, int max_outer,
                       femd_server *out);
int femd_server_submit(femd_server S, int64_t tag, femd_csr A, const double *b_dev, double *x_dev, int64_t *bytes);
int femd_server_step(femd_server S, int max_done, int64_t *tags, int *iterations, double *residuals, int *converged,
                     int *ndone);
int femd_server_status(femd_server S, int *queued, int *running, int *groups, int64_t *bytes_in_use);
void femd_server_free(femd_server S);

/* ---------------------------------------------------------------- time stepping
 * Stepping with a fixed matrix pair: u_{k+1} solves A u_{k+1} = M u_k (backward Euler for
 * M u' + K u = 0 with A = M + dt K).  No reference counterpart as one call: the reference writes
 * this loop in Python (example13.ipynb: b = M u, then a solve with A, each step); this replaces
 * that loop's SpMV and solve.  Each step: b = M u; up to `rounds` FP64 rounds r = b - A u, each
 * of which, unless ||r|| <= tol ||b||, runs a segment of at most max_inner iterations of the
 * scaled Jacobi-CG in `precision` (FEMD_FP32 or FEMD_FP64) on the correction, until its residual
 * has dropped by inner_tol (0 < inner_tol < 1) or its estimate of ||b - A u|| meets tol ||b||.
 * The first round of a step starts from no search direction (u_k is the warm start), later
 * rounds continue the same CG from the FP64 residual (reliable updates).  With use_graph the step
 * is captured once into a hipGraph and replayed nsteps times with no host synchronisation between
 * steps.  iters[k] (may be NULL) gets step k's inner iterations; info.iterations their sum,
 * info.breakdown the number of steps done, info.residual the largest ||b - A u|| / ||b|| that any
 * step ended with (each step checks its own FP64 residual after its rounds).  Each step gets the
 * verdict of one reference CG solve (krylov/cg.hpp:222-237): accepted when its residual is within
 * tol (1 + 1e-6) + 1e-15, or within 10 tol when its rounds stopped before using all
 * rounds x max_inner iterations; *failed (may be NULL) receives the steps not accepted and
 * info.converged says whether there were none. */
int femd_heat_run(femd_csr M, femd_csr A, double *u, int nsteps, double tol, int rounds, double inner_tol, int max_inner,
                  int precision, int use_graph, int *iters, int *failed, femd_solve_info *info);
int femd_heat_run_dev(femd_csr M, femd_csr A, double *u_dev, int nsteps, double tol, int rounds, double inner_tol,
                      int max_inner, int precision, int use_graph, int *iters, int *failed, femd_solve_info *info,
                      femd_stream s);

/* ---------------------------------------------------------------- timing
 * Kernel-time of the library's work on a stream, measured with HIP events around it. */
/* Sampled timing of the inner CG's product kernel (the refined solves, batches and the time
 * stepper): while enabled, every 8th launch outside graph capture is bracketed by HIP events on
 * its stream (solves then run without graphs); read returns the number of timed launches and
 * their total milliseconds.  femd_inner_timer resets the samples.  No counterpart in the
 * reference. */
int femd_inner_timer(int enable);
int femd_inner_timer_read(int64_t *launches, double *total_ms);
/* The launch shape (e.g. "tile256x2048") of the last timed product launch, NUL-terminated in
 * name[0 .. len); empty before the first.  No counterpart in the reference. */
int femd_inner_timer_shape(char *name, int len);

/* ---------------------------------------------------------------- value dictionaries
 * The inner CG (refined solves, batches, time stepping) reads its scaled matrix through a
 * lossless per-tile value dictionary whenever every SpMV tile holds at most 256 distinct values
 * (constant-coefficient problems on structured meshes): 1 byte per entry instead of 4 or 8, the
 * same values bit for bit, so results do not change.  mode 1 allows it (the default), 0 turns it
 * off, -1 only queries; returns the previous mode.  No counterpart in the reference. */
int femd_value_dictionary(int mode);
/* The inner CG (refined solves, batches, time stepping, the server) works on P A P^T, P numbering
 * every system's unknowns in the order its rows first meet them as columns, so that an SpMV tile's
 * rows and the columns it gathers lie close together; vectors at the interface stay in the
 * caller's order.  mode 1 (the default) uses that order, 0 the original one, -1 only queries;
 * returns the previous mode.  Results differ between the two only by the rounding of the inner
 * iterations.  No counterpart in the reference. */
int femd_inner_order(int mode);
/* How femd_assemble_matrix / femd_assemble_matrix_add form a dx group's element matrices: mode 0
 * (the default) with the exact kernels, in the reference's operation order
 * (forms/assembly_2d.hpp:102-126), bit-identical to it; mode 1 on the FP64 matrix cores, as one
 * GEMM over the cells, K_e[ij] = sum_k D_e[k] R[k][ij] with k running over (pair of reference
 * rows, quadrature point) -- the same sums in another order, within a few units of the last
 * place of the exact values (groups the kernel is not built for fall back to the exact kernels);
 * modes 2 and 3 the same GEMM with D and R rounded to f16 (2) or bf16 (3) and FP32 accumulation
 * (v_mfma_f32_16x16x16*), values within a few units of the 16-bit rounding -- for experiments with
 * low-precision local matrices; -1 only queries.  Returns the previous mode.  Boundary (ds) groups and vectors always use the
 * exact kernels.  No counterpart in the reference. */
int femd_assembly_mode(int mode);
/* Measurement: the average time (ms, HIP events on the library stream, `reps` launches after one
 * untimed) of the element-matrix kernel alone for a dx group -- mode 0 the exact kernel, 1 the
 * matrix-core one, 2 / 3 its f16 / bf16 variant -- and (mfma_flops, may be NULL) the matrix-core
 * flops one launch issues.  No
 * counterpart in the reference. */
int femd_assembly_kernel_time(femd_space V, int quad_degree, int nterms, const int *a, const int *b,
                              const femd_coefficient *coeffs, int mode, int reps, double *ms, double *mfma_flops);
/* The order femd_inner_order uses, computed on the host (no device needed): perm[k] (n) the
 * unknown at position k -- within every segment (row ranges segments[0..nseg], nseg 0: all rows),
 * the order a row-by-row scan first meets the unknowns as columns, unmet ones last -- and the
 * reordered pattern P A P^T: indptr2 (n + 1), indices2 (nnz, each row ascending), emap[p] (nnz)
 * the entry of the original pattern behind entry p.  FEMD_EINVAL when a segment's rows have
 * columns outside it.  No counterpart in the reference. */
int femd_inner_order_pattern(int n, const int *indptr, const int *indices, int nseg, const int64_t *segments, int *perm,
                             int *indptr2, int *indices2, int *emap);
/* The same order and reordered pattern computed on the GPU (host arrays in and out; the inner CG
 * builds it this way while femd_device_setup is on): first positions by an atomic minimum over the
 * entries, two radix sorts.  Equal to femd_inner_order_pattern's, entry for entry.  No counterpart
 * in the reference. */
int femd_inner_order_pattern_gpu(int n, const int *indptr, const int *indices, int nseg, const int64_t *segments, int *perm,
                                 int *indptr2, int *indices2, int *emap);
/* Where the set-up is built: mode 1 (the default) on the GPU -- a space's pattern and gather lists
 * (radix sorts of the element entries and cell dofs), the inner CG's first-met order and reordered
 * pattern, and the column lists of the SpMV tiles (one workgroup per tile sorting its columns in
 * LDS) -- 0 on the host (host.cpp's builders, which give the same arrays), -1 only queries;
 * returns the previous mode.  No counterpart in the reference. */
int femd_device_setup(int mode);
/* A space's pattern and its gather lists built afresh on the host (where 0, the builder femd_pattern
 * uses) or on the GPU (where 1, the one assembly uses while femd_device_setup is on): nnz, ncon (the
 * element entries landing in the pattern), indptr (dim + 1), indices (nnz), mptr (nnz + 1) and
 * mlist (ncon) -- stored entry p receives element entries mlist[mptr[p] .. mptr[p + 1]), element
 * entry (c nloc + a) nloc + b, ascending -- vptr (dim + 1) and vlist (the unconstrained
 * (cell, local dof) pairs, c nloc + a, ascending per dof).  Any array may be NULL (sizes only).  For
 * checking the two builders against each other.  No counterpart in the reference. */
int femd_pattern_lists(femd_space V, int where, int64_t *nnz, int64_t *ncon, int *indptr, int *indices, int *mptr,
                       int *mlist, int *vptr, int *vlist);
int femd_event_timer_start(femd_stream s);
int femd_event_timer_stop(femd_stream s, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* FEMD_HIP_H */
