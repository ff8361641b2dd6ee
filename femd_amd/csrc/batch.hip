// batch.hip -- many independent sparse SPD systems solved together (the batching pattern; the
// reference solves one system per call).
//
// The systems are concatenated into one block-diagonal CSR matrix held as an ordinary matrix
// handle whose pattern is cut into segments, one per system: SpMV tiles and the 4096-row vector
// chunks never straddle two systems, so every tile or chunk belongs to one system.  The inner CG
// engine (cg_engine.hpp) runs every system's Jacobi-preconditioned CG with its own scalars and
// reliable FP64 updates, in FP32 or FP64; kernels return at once for tiles and chunks of systems
// that are not running, so uneven sizes and iteration counts share one launch sequence, captured
// once into a hipGraph and replayed.
#include "batch.hpp"

#include <cmath>

using namespace femd_amd;

namespace femd_amd {

void batch_copy_values(femd_batch_s *B, hipStream_t s)
{
    for (int p = 0; p < B->np; ++p)
    {
        femd_csr_s *M = B->members[p];
        const int64_t nz = M->pat->nnz;
        if (nz) HIPC(hipMemcpyAsync(B->A->v.as<double>() + B->voff[p], M->v.p, 8 * size_t(nz), hipMemcpyDeviceToDevice, s));
    }
    B->A->v32_valid = false;
    B->A->v32s_valid = false;
    B->A->v64s_valid = false;
    scaled_values(B->A.get(), s, B->reorder);
}

std::unique_ptr<femd_batch_s> batch_build(int np, const femd_csr *mats, hipStream_t s, bool reorder)
{
    if (np < 1) fail(FEMD_EINVAL, "batch: at least one system");
    if (!mats) fail(FEMD_EINVAL, "batch: null argument");
    auto B = std::make_unique<femd_batch_s>();
    B->np = np;
    B->reorder = reorder;
    B->members.assign(mats, mats + np);
    B->roff.assign(np + 1, 0);
    B->voff.assign(np + 1, 0);
    for (int p = 0; p < np; ++p)
    {
        femd_csr_s *M = check_csr(mats[p]);
        if (M->pat->n != M->pat->m) fail(FEMD_EINVAL, "batch: every system must be square");
        B->roff[p + 1] = B->roff[p] + M->pat->n;
        B->voff[p + 1] = B->voff[p] + M->pat->nnz;
        B->max_rows = std::max<int64_t>(B->max_rows, M->pat->n);
    }
    B->nrows = B->roff[np];
    B->nnz = B->voff[np];
    if (B->nnz >= (int64_t(1) << 31) || B->nrows >= (int64_t(1) << 31)) fail(FEMD_EINVAL, "batch: more than 2^31 rows or entries");
    // the block-diagonal pattern (host copy for the tile plans), one segment per system
    std::vector<int> ip(size_t(B->nrows) + 1), ix(size_t(B->nnz));
    for (int p = 0; p < np; ++p)
    {
        const DevPattern &P = *mats[p]->pat;
        const int r0 = static_cast<int>(B->roff[p]), v0 = static_cast<int>(B->voff[p]);
        for (int i = 0; i <= P.n; ++i) ip[size_t(r0 + i)] = v0 + P.h_ip[size_t(i)];
        for (int64_t k = 0; k < P.nnz; ++k) ix[size_t(v0 + k)] = r0 + P.h_ix[size_t(k)];
    }
    B->A = std::make_unique<femd_csr_s>();
    B->A->pat = make_pattern(static_cast<int>(B->nrows), static_cast<int>(B->nrows), std::move(ip), std::move(ix), s);
    B->A->pat->segments = B->roff;
    B->A->v = DevBuf(8 * size_t(std::max<int64_t>(B->nnz, 1)));
    batch_copy_values(B.get(), s);
    HIPC(hipStreamSynchronize(s));
    return B;
}

} // namespace femd_amd

namespace {

void batch_solve(femd_batch_s *B, const double *b, double *x, double tol, double inner_tol, int max_outer, int precision,
                 int use_graph, int *iters, double *res, femd_solve_info *info, hipStream_t s)
{
    std::memset(info, 0, sizeof(*info));
    if (precision != FEMD_FP32 && precision != FEMD_FP64) fail(FEMD_EINVAL, "batch: precision is FEMD_FP32 or FEMD_FP64");
    if (!(tol >= 0.0) || !(inner_tol > 0.0 && inner_tol < 1.0) || max_outer < 1)
        fail(FEMD_EINVAL, "batch: need tol >= 0, 0 < inner_tol < 1, max_outer >= 1");
    CGEngine &E = B->engine[precision == FEMD_FP32 ? 0 : 1];
    E.setup(B->A.get(), precision, s, B->reorder);
    const CGEngine::Result r = E.solve(b, x, tol, inner_tol, max_outer, use_graph != 0, iters, res, s);
    info->converged = r.converged;
    info->iterations = r.iterations;
    info->breakdown = r.rounds;
    info->residual = r.worst;
    info->precond_residual = r.worst;
}

} // namespace

extern "C" {

int femd_batch_create(int np, const femd_csr *mats, femd_batch *out)
{
    return api([&] {
        if (!out) fail(FEMD_EINVAL, "batch: null argument");
        *out = batch_build(np, mats, S(nullptr)).release();
    });
}

int femd_batch_info(femd_batch B, int64_t *nrows, int64_t *nnz, int64_t *row_offsets)
{
    return api([&] {
        if (!B) fail(FEMD_EINVAL, "null batch");
        if (nrows) *nrows = B->nrows;
        if (nnz) *nnz = B->nnz;
        if (row_offsets) std::copy(B->roff.begin(), B->roff.end(), row_offsets);
    });
}

int femd_batch_update_values(femd_batch B)
{
    return api([&] {
        if (!B) fail(FEMD_EINVAL, "null batch");
        batch_copy_values(B, S(nullptr));
    });
}

int femd_batch_solve_dev(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer, int precision,
                         int use_graph, int *iterations, double *residuals, femd_solve_info *info, femd_stream st)
{
    return api([&] {
        if (!B || !info) fail(FEMD_EINVAL, "batch: null argument");
        batch_solve(B, b, x, tol, inner_tol, max_outer, precision, use_graph, iterations, residuals, info, S(st));
    });
}

int femd_batch_solve(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer, int precision,
                     int use_graph, int *iterations, double *residuals, femd_solve_info *info)
{
    return api([&] {
        if (!B || !info) fail(FEMD_EINVAL, "batch: null argument");
        hipStream_t s = S(nullptr);
        DevBuf bd = upload(b, size_t(B->nrows), s), xd = upload(x, size_t(B->nrows), s);
        batch_solve(B, bd.as<double>(), xd.as<double>(), tol, inner_tol, max_outer, precision, use_graph, iterations, residuals,
                   info, s);
        if (B->nrows) HIPC(hipMemcpyAsync(x, xd.p, 8 * size_t(B->nrows), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_batch_solve_refined_dev(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer,
                                 int use_graph, int *iterations, double *residuals, femd_solve_info *info, femd_stream st)
{
    return femd_batch_solve_dev(B, b, x, tol, inner_tol, max_outer, FEMD_FP32, use_graph, iterations, residuals, info, st);
}

int femd_batch_solve_refined(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer,
                             int use_graph, int *iterations, double *residuals, femd_solve_info *info)
{
    return femd_batch_solve(B, b, x, tol, inner_tol, max_outer, FEMD_FP32, use_graph, iterations, residuals, info);
}

int femd_batch_dictionary(femd_batch B, int precision, int64_t *distinct, int *max_per_tile)
{
    return api([&] {
        if (!B) fail(FEMD_EINVAL, "null batch");
        if (precision != FEMD_FP32 && precision != FEMD_FP64) fail(FEMD_EINVAL, "batch: precision is FEMD_FP32 or FEMD_FP64");
        const CGEngine &E = B->engine[precision == FEMD_FP32 ? 0 : 1];
        const ValueDict &d = precision == FEMD_FP32 ? B->A->d32s : B->A->d64s;
        if (distinct) *distinct = E.use_dict ? d.distinct : -1;
        if (max_per_tile) *max_per_tile = d.valid ? d.max_distinct : -1;
    });
}

void femd_batch_free(femd_batch B)
{
    if (B && g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
    delete B;
}

} // extern "C"
