// kernels.hpp -- launchers of the gfx950 kernels (kernels.hip).  All take a stream and enqueue.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace femd_amd {

constexpr int CHUNK = 4096;                // the reference's reduction chunk (util/omp.hpp:91)
constexpr int MAX_TERMS = 8;
constexpr int MAX_REGS = 64;

constexpr int MAX_INPUTS = 8;

// The known fields a coefficient program reads (INPUT a: p[a][cell * nq + q]).
struct InputSet {
    const double *p[MAX_INPUTS];
};

// Coefficients of up to MAX_TERMS terms: a device array [cell * nq + q] or one constant.
struct TermSet {
    int nt;
    int a[MAX_TERMS], b[MAX_TERMS];
    const double *c[MAX_TERMS];
    double cval[MAX_TERMS];
};

// ---- assembly -------------------------------------------------------------------------------
// false: no kernel for this local size, or the reference tables do not fit LDS
bool launch_elem_matrix(int nv, int nloc, int nc, int nq, const double *geom, const double *tab, const double *qw,
                        const double *qxi, const double *qeta, const TermSet &T, double *out, hipStream_t s);
bool launch_elem_vector(int nv, int nloc, int nc, int nq, const double *geom, const double *tab, const double *qw,
                        const double *qxi, const double *qeta, const TermSet &T, double *out, hipStream_t s);
// The matrix-core element kernel (mfma.hip): Ke[ij] = sum_k D_e[k] R[k][ij].  Step s = c spc + j of
// the K loop holds combo c's points q = 4 j + 0..3 (points past nq padded with zero rows of R):
// combo c pairs reference row alpha_c (bits 2c of alpha) on the test side with beta_c on the trial
// side.  rt is R, (4 ks) x (16 ntile) doubles, ks = ncombo spc; ocol[n] the entries (i N + j)
// computed column n goes to (.y = -1: none).
constexpr int MFMA_SPCMAX = 8;               // steps per combo: rules of at most 32 points
constexpr int MFMA_UMAX = 4;                 // distinct coefficient arrays per group
constexpr size_t MFMA_LDS_MAX = 65536;       // R is staged in LDS up to this size
struct MfmaJob {
    int nc, nq, N, ncombo, spc, ks, ntile, nout;
    unsigned alpha, beta;
    const double *geom, *qw, *qxi, *qeta, *rt;
    const int2 *ocol;
    double *E;
};
// The terms as the kernel reads them: codes, the constant of a constant term (slot -1), or the
// slot of its coefficient array among the nu distinct ones.
struct MfmaTerms {
    int nt, nu;
    int a[MAX_TERMS], b[MAX_TERMS], slot[MAX_TERMS];
    double cval[MAX_TERMS];
    const double *cu[MFMA_UMAX];
};
size_t mfma_table_bytes(int ks, int ntile);
// false: the rule, the coefficient count, the tile count or the LDS size is outside what the kernel
// is built for (the exact kernel runs instead)
bool launch_elem_matrix_mfma(int nv, const MfmaJob &J, const TermSet &T, int nt, bool rlds, hipStream_t s);
// The same with D and R rounded to 16 bits (f16, or bf16 with bf16 set) and FP32 accumulation
// (v_mfma_f32_16x16x16*): J.rt holds R in 16 bits laid out [step of 16 k][entry][16 k], J.spc the
// steps of 16 per combo (at most two).  For the precision experiment.
bool launch_elem_matrix_mfma16(int nv, const MfmaJob &J, const TermSet &T, int nt, bool bf16, hipStream_t s);

void launch_gather(int64_t n, const int *ptr, const int *list, const double *src, double *dst, bool plus_zero, hipStream_t s);
// How a gather slot combines its element-order sum with the destination (k_gather).
enum { GATHER_SET = 0, GATHER_SET_PZ = 1, GATHER_ADD = 2, GATHER_CONTINUE = 3 };
void launch_gather_mode(int64_t n, const int *pos, const int *ptr, const int *list, const double *src, double *dst, int mode,
                        hipStream_t s);
// Boundary-facet element matrices / vectors (false: no kernel for this local size).
bool launch_facet_matrix(int nloc, int ne, int nq, const int *var, const double *w, const double *Ji, const double *tab,
                         const TermSet &T, double *out, hipStream_t s);
bool launch_facet_vector(int nloc, int ne, int nq, const int *var, const double *w, const double *Ji, const double *tab,
                         const TermSet &T, double *out, hipStream_t s);
// A coefficient program at np given points with their normals.
void launch_coefficient_pts(int64_t np, const double *x, const double *y, const double *nx, const double *ny, int ninstr,
                            const int *code, const double *consts, double *out, hipStream_t s);
void launch_coefficient(int nc, int nq, int nv, const double *geom, const double *qxi, const double *qeta, int ninstr,
                        const int *code, int nreg, const double *consts, const InputSet &in, double *out, hipStream_t s);
// A known field (coefficients u, the cell_dof of its space, -1 eliminated) at the nq points of every
// cell: its value (code 0) or x / y derivative (1 / 2), Cache2D::eval_multi's sums
// (fe/cache_2d.hpp:245-288).  tab: the field space's reference table at the points.  false: no
// kernel for this local size.
bool launch_eval_field(int nv, int nloc, int nc, int nq, const int *cell_dof, const double *u, const double *tab,
                       const double *geom, const double *qxi, const double *qeta, int code, double *out, hipStream_t s);

// ---- SpMV -----------------------------------------------------------------------------------

// ---- ordered reductions (util/omp.hpp:135-192) -----------------------------------------------
int nchunks(int64_t n);
void launch_dot_parts(int64_t n, const double *x, const double *y, double *parts, const int *stop, hipStream_t s);
void launch_sum_parts(int np, const double *parts, double *out, hipStream_t s);

// ---- vector kernels ---------------------------------------------------------------------------
void launch_sub(int64_t n, const double *b, const double *a, double *r, hipStream_t s);             // r = b - a
void launch_mul(int64_t n, const double *a, const double *b, double *c, hipStream_t s);             // c = a * b
void launch_copy(int64_t n, const double *a, double *b, hipStream_t s);
void launch_fill(int64_t n, double v, double *a, hipStream_t s);
void launch_gather_perm(int n, const int *perm, const double *src, double *dst, hipStream_t s);     // dst[k] = src[perm[k]]
void launch_scatter_perm(int n, const int *perm, const double *src, double *dst, hipStream_t s);    // dst[perm[k]] = src[k]
void launch_inv_diag(int n, const int *diag, const double *v, double *inv, int *bad_row, hipStream_t s);
void launch_diag_positions(int n, const int *ip, const int *ix, int *diag, hipStream_t s);
void launch_take(int64_t n, const int *idx, const double *src, double *dst, hipStream_t s);         // dst[i] = src[idx[i]]
void launch_to_float(int64_t n, const double *a, float *b, hipStream_t s);
void launch_transpose_positions(int n, const int *ip, const int *ix, int *tpos, hipStream_t s);
void launch_asymmetry(int64_t nnz, const int *tpos, const double *v, double *out2, hipStream_t s);

// ---- CG (krylov/cg.hpp) -------------------------------------------------------------------------
struct CGState {
    double nb, tolnb, rz, rr, pAp, alpha, beta, rn;
    int it, max_iter, stop, breakdown, converged_break, pad;
};
void launch_cg_alpha(const double *parts, int np, CGState *st, hipStream_t s);
void launch_cg_update(int64_t n, double *x, double *r, const double *p, const double *Ap, const double *inv,
                      double *parts_rr, double *parts_rz, CGState *st, hipStream_t s);
void launch_cg_beta(const double *parts_rr, const double *parts_rz, int np, int jacobi_fused, CGState *st, hipStream_t s);
void launch_cg_rz(const double *parts, int np, CGState *st, hipStream_t s);
void launch_cg_pupdate(int64_t n, double *p, const double *r, const double *inv, const CGState *st, hipStream_t s);

// ---- triangular sweeps and factorization (sparse/precond.hpp), one workgroup, level by level ----
void launch_ilu0_factor(int nlev, const int *lptr, const int *lrows, const int *ip, const int *ix, const int *diag,
                        double *v, int *bad_row, hipStream_t s);
// the level-parallel factor (one launch per level, one wave per row) and whether a pattern fits it
// (host arrays); the smallest pivot into out (one double on the device)
bool ilu0_levels_fit(int n, const int *ip, const int *ix, const int *diag);
void launch_ilu0_levels(int nlev, const int *lptr_host, const int *lrows, const int *ip, const int *ix, const int *diag,
                        double *v, int *bad_row, hipStream_t s);
void launch_min_pivot(int n, const int *diag, const double *v, double *out, hipStream_t s);
void launch_ilu0_apply(int n, int nlevL, const int *lptrL, const int *lrowsL, int nlevU, const int *lptrU,
                       const int *lrowsU, const int *ip, const int *ix, const int *diag, const double *v,
                       const double *r, double *z, const int *stop, hipStream_t s);
void launch_ssor_apply(int n, int nlevL, const int *lptrL, const int *lrowsL, int nlevU, const int *lptrU,
                       const int *lrowsU, const int *ip, const int *ix, const int *diag, const double *v, double omega,
                       const double *r, double *z, const int *stop, hipStream_t s);

// ---- the planned one-workgroup sweeps (sweep.hip, sweep.hpp) --------------------------------------
void launch_sweep(int kind, int ereg, int nsteps, const void *stream, const void *first, uint32_t zunit, double omega,
                  double *out, const int *stop, hipStream_t s);
void launch_sweep_fill(int n, const int *in_pos, const int *perm, const double *src, double *stream, const int *stop,
                       hipStream_t s);
void launch_sweep_scatter(int n, const int *to, const double *src, double *dst, const int *stop, hipStream_t s);
void launch_sweep_pack(int64_t nnz, const int *fmap, const int *bmap, const double *v, double *fwd, double *bwd, hipStream_t s);

// ---- fast (non-ordered) reductions and FP32 CG for the mixed-precision path ----------------------
int fast_blocks(int64_t n);
void launch_fdot(int64_t n, const double *x, const double *y, double *parts, hipStream_t s);
void launch_fsum(int np, const double *parts, double *out, hipStream_t s);
void launch_axpy_dev(int64_t n, const double *alpha, double sign, const double *x, double *y, hipStream_t s);
void launch_scale_dev(int64_t n, const double *num, const double *den, double *x, hipStream_t s);   // x *= num/den
void launch_axpy(int64_t n, double alpha, const double *x, double *y, hipStream_t s);
void launch_scale(int64_t n, double alpha, double *x, hipStream_t s);
void launch_lincomb(int64_t n, int k, const double *y, const double *const *V, double *x, hipStream_t s);

} // namespace femd_amd
