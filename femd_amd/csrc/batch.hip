// batch.hip -- many independent sparse SPD systems solved together (the batching pattern; the
// reference solves one system per call).
//
// The systems are concatenated into one block-diagonal CSR matrix.  SpMV row blocks and the
// 4096-row reduction chunks never straddle two systems, so every block or chunk belongs to one
// system (bpid / cpid), and every dot product is a per-system (segmented) reduction.  Each system
// has its own CG scalars and a done flag on the device; a finished system's blocks return at once,
// so uneven sizes and iteration counts share one launch sequence.  The inner FP32 Jacobi-PCG
// iteration (six kernels) is captured once into a hipGraph and replayed in batches between two
// read-backs of the number of unfinished systems.  FP64 refinement wraps it per system:
// r = b - A x in FP64, the FP32 solve of A d = r / ||r||, x += ||r|| d.
#include "internal.hpp"

#include <cmath>
#include <numeric>

using namespace femd_amd;

namespace femd_amd {
namespace {

constexpr int BT = 256;
constexpr int BCHUNK = 4096;

struct BState {
    double rz, alpha, beta, rr, thr;
    int it, done, max_iter, breakdown;
};

__device__ inline double bsum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = BT / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double out = sh[0];
    __syncthreads();
    return out;
}

// CSR-stream SpMV over the block-diagonal matrix; a block of a finished system does nothing.
template <typename T>
__global__ __launch_bounds__(SPMV_THREADS) void k_bspmv(const int *__restrict__ blocks, const int *__restrict__ bpid,
                                                        const int *__restrict__ ip, const int *__restrict__ ix,
                                                        const T *__restrict__ v, const T *__restrict__ x, T *__restrict__ y,
                                                        const BState *st)
{
    if (st && st[bpid[blockIdx.x]].done) return;
    __shared__ T prod[SPMV_CAP64];
    const int r0 = blocks[blockIdx.x], r1 = blocks[blockIdx.x + 1];
    const int p0 = ip[r0], len = ip[r1] - p0;
    if (len > SPMV_CAP64)
    {
        if (threadIdx.x == 0)
        {
            T s = 0;
            for (int p = p0; p < p0 + len; ++p) s += v[p] * x[ix[p]];
            y[r0] = s;
        }
        return;
    }
    for (int k = threadIdx.x; k < len; k += blockDim.x) prod[k] = v[p0 + k] * x[ix[p0 + k]];
    __syncthreads();
    const int r = r0 + threadIdx.x;
    if (r < r1)
    {
        T s = 0;
        for (int k = ip[r] - p0, e = ip[r + 1] - p0; k < e; ++k) s += prod[k];
        y[r] = s;
    }
}

// per chunk: sum x_i y_i (FP64 accumulation) of the chunk's rows
template <typename T>
__global__ __launch_bounds__(BT) void k_bdot(const int64_t *__restrict__ clo, const int *__restrict__ cpid,
                                             const T *__restrict__ x, const T *__restrict__ y, double *__restrict__ parts,
                                             const BState *st)
{
    __shared__ double sh[BT];
    const int c = blockIdx.x;
    if (st && st[cpid[c]].done) return;
    double s = 0.0;
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT) s += static_cast<double>(x[i]) * static_cast<double>(y[i]);
    s = bsum(s, sh);
    if (threadIdx.x == 0) parts[c] = s;
}

// one thread per system: alpha = rz / p^T A p
__global__ void k_balpha(int np, const int *__restrict__ cptr, const double *__restrict__ parts, BState *st)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || st[p].done) return;
    double s = 0.0;
    for (int c = cptr[p]; c < cptr[p + 1]; ++c) s += parts[c];
    if (!(s > 0.0)) { st[p].breakdown = 1; st[p].done = 1; return; }
    st[p].alpha = st[p].rz / s;
}

__global__ __launch_bounds__(BT) void k_bupdate32(const int64_t *__restrict__ clo, const int *__restrict__ cpid, float *x,
                                                  float *r, const float *__restrict__ p, const float *__restrict__ Ap,
                                                  const float *__restrict__ inv, double *__restrict__ prr,
                                                  double *__restrict__ prz, const BState *st)
{
    __shared__ double sh[BT];
    const int c = blockIdx.x;
    const BState &S = st[cpid[c]];
    if (S.done) return;
    const float a = static_cast<float>(S.alpha);
    double srr = 0.0, srz = 0.0;
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT)
    {
        x[i] = fmaf(a, p[i], x[i]);
        const float ri = fmaf(-a, Ap[i], r[i]);
        r[i] = ri;
        const double e = static_cast<double>(ri) * ri;
        srr += e;
        srz += e * inv[i];
    }
    srr = bsum(srr, sh);
    srz = bsum(srz, sh);
    if (threadIdx.x == 0) { prr[c] = srr; prz[c] = srz; }
}

// one thread per system: the stopping test, beta, the iteration cap
__global__ void k_bbeta(int np, const int *__restrict__ cptr, const double *__restrict__ prr, const double *__restrict__ prz,
                        BState *st)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || st[p].done) return;
    double a = 0.0, b = 0.0;
    for (int c = cptr[p]; c < cptr[p + 1]; ++c) { a += prr[c]; b += prz[c]; }
    BState &S = st[p];
    S.rr = a;
    S.it += 1;
    if (a <= S.thr) { S.done = 1; return; }
    S.beta = b / S.rz;
    S.rz = b;
    if (S.it >= S.max_iter) S.done = 1;
    else if (!(b > 0.0)) { S.breakdown = 1; S.done = 1; }
}

__global__ __launch_bounds__(BT) void k_bpupdate32(const int64_t *__restrict__ clo, const int *__restrict__ cpid, float *p,
                                                   const float *__restrict__ r, const float *__restrict__ inv, const BState *st)
{
    const int c = blockIdx.x;
    const BState &S = st[cpid[c]];
    if (S.done) return;
    const float beta = static_cast<float>(S.beta);
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT) p[i] = fmaf(beta, p[i], inv[i] * r[i]);
}

// start of an inner solve: r = r64 / ||r64|| (per system), x = 0, p = D^{-1} r, parts of r^T p
__global__ __launch_bounds__(BT) void k_binit32(const int64_t *__restrict__ clo, const int *__restrict__ cpid,
                                                const double *__restrict__ r64, const double *__restrict__ scale, float *r,
                                                float *x, float *p, const float *__restrict__ inv, double *__restrict__ parts)
{
    __shared__ double sh[BT];
    const int c = blockIdx.x;
    const double sc = scale[cpid[c]];
    double s = 0.0;
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT)
    {
        const float ri = static_cast<float>(r64[i] * sc);
        const float pi = inv[i] * ri;
        r[i] = ri; x[i] = 0.0f; p[i] = pi;
        s += static_cast<double>(ri) * pi;
    }
    s = bsum(s, sh);
    if (threadIdx.x == 0) parts[c] = s;
}

// one thread per system: rz, the threshold and the cap of the inner solve; inactive systems are done
__global__ void k_bstart(int np, const int *__restrict__ cptr, const double *__restrict__ parts, const int *__restrict__ active,
                         double thr, int max_iter, BState *st)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double s = 0.0;
    for (int c = cptr[p]; c < cptr[p + 1]; ++c) s += parts[c];
    BState &S = st[p];
    S.rz = s; S.it = 0; S.breakdown = 0; S.thr = thr; S.max_iter = max_iter; S.alpha = 0.0; S.beta = 0.0; S.rr = 0.0;
    S.done = active[p] ? 0 : 1;
    if (!S.done && !(s > 0.0)) S.done = 1;
}

__global__ __launch_bounds__(BT) void k_badd32(const int64_t *__restrict__ clo, const int *__restrict__ cpid,
                                               const double *__restrict__ back, const int *__restrict__ active,
                                               const float *__restrict__ d, double *x)
{
    const int c = blockIdx.x;
    const int pid = cpid[c];
    if (!active[pid]) return;
    const double sc = back[pid];
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT) x[i] += sc * static_cast<double>(d[i]);
}

__global__ void k_bcount(int np, const BState *st, int *out)
{
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int p = threadIdx.x; p < np; p += blockDim.x) mine += st[p].done ? 0 : 1;
    atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) out[0] = cnt;
}

__global__ void k_bsum_parts(int np, const int *__restrict__ cptr, const double *__restrict__ parts, double *out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double s = 0.0;
    for (int c = cptr[p]; c < cptr[p + 1]; ++c) s += parts[c];
    out[p] = s;
}

__global__ void k_offset_add(int64_t n, const int *src, int off, int *dst)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + off;
}

__global__ void k_inv32_diag(int64_t n, const int *__restrict__ diag, const double *__restrict__ v, float *inv, int *bad)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = diag[i];
    if (d < 0 || v[d] == 0.0) { atomicMin(bad, static_cast<int>(i)); inv[i] = 0.0f; return; }
    inv[i] = static_cast<float>(1.0 / v[d]);
}

inline int grid1(int64_t n) { return static_cast<int>(std::max<int64_t>(1, (n + BT - 1) / BT)); }

} // namespace
} // namespace femd_amd

struct femd_batch_s {
    int np = 0;
    int64_t nrows = 0, nnz = 0;
    std::vector<femd_csr> members;
    std::vector<int64_t> roff;                 // row offsets (np + 1)
    std::vector<int64_t> voff;                 // value offsets (np + 1)
    std::vector<int> h_cptr;
    DevBuf ip, ix, v, v32, blocks, bpid, clo, cpid, cptr, diag, inv32;
    int nblocks = 0, nch = 0;
    // solver work space
    DevBuf st, r32, x32, p32, Ap32, r64, t64, prr, prz, scal, active, back, cnt;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    int graph_iters = 0;
    ~femd_batch_s()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

void batch_copy_values(femd_batch_s *B, hipStream_t s)
{
    for (int p = 0; p < B->np; ++p)
    {
        femd_csr_s *A = B->members[p];
        const int64_t nz = A->pat->nnz;
        if (nz) HIPC(hipMemcpyAsync(B->v.as<double>() + B->voff[p], A->v.p, 8 * size_t(nz), hipMemcpyDeviceToDevice, s));
    }
    launch_to_float(B->nnz, B->v.as<double>(), B->v32.as<float>(), s);
    DevBuf bad(sizeof(int));
    const int big = 0x7fffffff;
    HIPC(hipMemcpyAsync(bad.p, &big, sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_inv32_diag, dim3(grid1(B->nrows)), dim3(BT), 0, s, B->nrows, B->diag.as<int>(), B->v.as<double>(),
                       B->inv32.as<float>(), bad.as<int>());
    int h = 0;
    HIPC(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (h != big) fail(FEMD_EZEROPIVOT, "batch: zero diagonal entry in row " + std::to_string(h) + " of the concatenated system");
}

// The six kernels of one inner FP32 Jacobi-PCG iteration over every unfinished system.
void inner_iteration(femd_batch_s *B, hipStream_t s)
{
    BState *st = B->st.as<BState>();
    const int np = B->np;
    hipLaunchKernelGGL(k_bspmv<float>, dim3(B->nblocks), dim3(SPMV_THREADS), 0, s, B->blocks.as<int>(), B->bpid.as<int>(),
                       B->ip.as<int>(), B->ix.as<int>(), B->v32.as<float>(), B->p32.as<float>(), B->Ap32.as<float>(), st);
    hipLaunchKernelGGL(k_bdot<float>, dim3(B->nch), dim3(BT), 0, s, B->clo.as<int64_t>(), B->cpid.as<int>(),
                       B->p32.as<float>(), B->Ap32.as<float>(), B->prr.as<double>(), st);
    hipLaunchKernelGGL(k_balpha, dim3(grid1(np)), dim3(BT), 0, s, np, B->cptr.as<int>(), B->prr.as<double>(), st);
    hipLaunchKernelGGL(k_bupdate32, dim3(B->nch), dim3(BT), 0, s, B->clo.as<int64_t>(), B->cpid.as<int>(), B->x32.as<float>(),
                       B->r32.as<float>(), B->p32.as<float>(), B->Ap32.as<float>(), B->inv32.as<float>(), B->prr.as<double>(),
                       B->prz.as<double>(), st);
    hipLaunchKernelGGL(k_bbeta, dim3(grid1(np)), dim3(BT), 0, s, np, B->cptr.as<int>(), B->prr.as<double>(), B->prz.as<double>(), st);
    hipLaunchKernelGGL(k_bpupdate32, dim3(B->nch), dim3(BT), 0, s, B->clo.as<int64_t>(), B->cpid.as<int>(), B->p32.as<float>(),
                       B->r32.as<float>(), B->inv32.as<float>(), st);
}

void launch_iterations(femd_batch_s *B, int iters, bool use_graph, hipStream_t s)
{
    if (!use_graph)
    {
        for (int k = 0; k < iters; ++k) inner_iteration(B, s);
        return;
    }
    if (!B->exec || B->graph_stream != s || B->graph_iters != iters)
    {
        if (B->exec) { HIPC(hipGraphExecDestroy(B->exec)); B->exec = nullptr; }
        if (B->graph) { HIPC(hipGraphDestroy(B->graph)); B->graph = nullptr; }
        HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < iters; ++k) inner_iteration(B, s);
        HIPC(hipStreamEndCapture(s, &B->graph));
        HIPC(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        B->graph_stream = s;
        B->graph_iters = iters;
    }
    HIPC(hipGraphLaunch(B->exec, s));
}

// Per-system ||b - A x||_2 (FP64) into host; r64 keeps the residual.
std::vector<double> residual_norms(femd_batch_s *B, const double *b, const double *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_bspmv<double>, dim3(B->nblocks), dim3(SPMV_THREADS), 0, s, B->blocks.as<int>(), B->bpid.as<int>(),
                       B->ip.as<int>(), B->ix.as<int>(), B->v.as<double>(), x, B->t64.as<double>(), static_cast<const BState *>(nullptr));
    launch_sub(B->nrows, b, B->t64.as<double>(), B->r64.as<double>(), s);
    hipLaunchKernelGGL(k_bdot<double>, dim3(B->nch), dim3(BT), 0, s, B->clo.as<int64_t>(), B->cpid.as<int>(),
                       B->r64.as<double>(), B->r64.as<double>(), B->prr.as<double>(), static_cast<const BState *>(nullptr));
    hipLaunchKernelGGL(k_bsum_parts, dim3(grid1(B->np)), dim3(BT), 0, s, B->np, B->cptr.as<int>(), B->prr.as<double>(), B->scal.as<double>());
    std::vector<double> h(B->np);
    HIPC(hipMemcpyAsync(h.data(), B->scal.p, 8 * size_t(B->np), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    for (double &v : h) v = std::sqrt(v);
    return h;
}

std::vector<double> vector_norms(femd_batch_s *B, const double *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_bdot<double>, dim3(B->nch), dim3(BT), 0, s, B->clo.as<int64_t>(), B->cpid.as<int>(), b, b,
                       B->prr.as<double>(), static_cast<const BState *>(nullptr));
    hipLaunchKernelGGL(k_bsum_parts, dim3(grid1(B->np)), dim3(BT), 0, s, B->np, B->cptr.as<int>(), B->prr.as<double>(), B->scal.as<double>());
    std::vector<double> h(B->np);
    HIPC(hipMemcpyAsync(h.data(), B->scal.p, 8 * size_t(B->np), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    for (double &v : h) v = std::sqrt(v);
    return h;
}

void batch_solve(femd_batch_s *B, const double *b, double *x, double tol, double inner_tol, int max_outer, int use_graph,
                 int *iters, double *res, femd_solve_info *info, hipStream_t s)
{
    const int np = B->np;
    std::memset(info, 0, sizeof(*info));
    std::vector<double> nb = vector_norms(B, b, s);
    std::vector<int> act(np), inner(np, 0), conv(np, 0);
    std::vector<double> back(np), scale(np), rel(np, 0.0);
    int outer = 0;
    Pinned hc(sizeof(int));
    for (;;)
    {
        std::vector<double> rn = residual_norms(B, b, x, s);
        int nact = 0;
        for (int p = 0; p < np; ++p)
        {
            rel[p] = nb[p] > 0.0 ? rn[p] / nb[p] : 0.0;
            conv[p] = rel[p] <= tol;
            act[p] = !conv[p] && outer < max_outer;
            nact += act[p];
            back[p] = rn[p];
            scale[p] = rn[p] > 0.0 ? 1.0 / rn[p] : 0.0;
        }
        if (nact == 0) break;
        ++outer;
        HIPC(hipMemcpyAsync(B->active.p, act.data(), 4 * size_t(np), hipMemcpyHostToDevice, s));
        HIPC(hipMemcpyAsync(B->back.p, back.data(), 8 * size_t(np), hipMemcpyHostToDevice, s));
        HIPC(hipMemcpyAsync(B->scal.p, scale.data(), 8 * size_t(np), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_binit32, dim3(B->nch), dim3(BT), 0, s, B->clo.as<int64_t>(), B->cpid.as<int>(), B->r64.as<double>(),
                           B->scal.as<double>(), B->r32.as<float>(), B->x32.as<float>(), B->p32.as<float>(), B->inv32.as<float>(),
                           B->prr.as<double>());
        int maxn = 0;
        for (int p = 0; p < np; ++p) maxn = std::max<int64_t>(maxn, B->roff[p + 1] - B->roff[p]);
        hipLaunchKernelGGL(k_bstart, dim3(grid1(np)), dim3(BT), 0, s, np, B->cptr.as<int>(), B->prr.as<double>(),
                           B->active.as<int>(), inner_tol * inner_tol, std::max(10, 10 * maxn), B->st.as<BState>());
        int left = nact, batch = 32;
        while (left > 0)
        {
            launch_iterations(B, batch, use_graph != 0, s);
            hipLaunchKernelGGL(k_bcount, dim3(1), dim3(BT), 0, s, np, B->st.as<BState>(), B->cnt.as<int>());
            HIPC(hipMemcpyAsync(hc.p, B->cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPC(hipStreamSynchronize(s));
            left = *hc.as<int>();
        }
        std::vector<BState> hs(np);
        HIPC(hipMemcpyAsync(hs.data(), B->st.p, sizeof(BState) * np, hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_badd32, dim3(B->nch), dim3(BT), 0, s, B->clo.as<int64_t>(), B->cpid.as<int>(), B->back.as<double>(),
                           B->active.as<int>(), B->x32.as<float>(), x);
        HIPC(hipStreamSynchronize(s));
        for (int p = 0; p < np; ++p) if (act[p]) inner[p] += hs[p].it;
    }
    int nconv = 0;
    double worst = 0.0;
    for (int p = 0; p < np; ++p)
    {
        if (iters) iters[p] = inner[p];
        if (res) res[p] = rel[p];
        nconv += conv[p];
        worst = std::max(worst, rel[p]);
        info->iterations += inner[p];
    }
    info->converged = nconv == np;
    info->breakdown = outer;
    info->residual = worst;
    info->precond_residual = worst;
}

} // namespace

extern "C" {

int femd_batch_create(int np, const femd_csr *mats, femd_batch *out)
{
    return api([&] {
        if (np < 1) fail(FEMD_EINVAL, "batch: at least one system");
        hipStream_t s = S(nullptr);
        auto B = std::make_unique<femd_batch_s>();
        B->np = np;
        B->members.assign(mats, mats + np);
        B->roff.assign(np + 1, 0);
        B->voff.assign(np + 1, 0);
        for (int p = 0; p < np; ++p)
        {
            femd_csr_s *A = check_csr(mats[p]);
            if (A->pat->n != A->pat->m) fail(FEMD_EINVAL, "batch: every system must be square");
            B->roff[p + 1] = B->roff[p] + A->pat->n;
            B->voff[p + 1] = B->voff[p] + A->pat->nnz;
        }
        B->nrows = B->roff[np];
        B->nnz = B->voff[np];
        if (B->nnz >= (int64_t(1) << 31) || B->nrows >= (int64_t(1) << 31)) fail(FEMD_EINVAL, "batch: more than 2^31 rows or entries");
        // concatenated pattern: indptr on the host, indices shifted on the device
        std::vector<int> ip(B->nrows + 1);
        std::vector<int> blocks{0}, bpid;
        std::vector<int64_t> clo{0};
        std::vector<int> cpid, cptr{0};
        for (int p = 0; p < np; ++p)
        {
            const DevPattern &P = *mats[p]->pat;
            const int r0 = static_cast<int>(B->roff[p]), v0 = static_cast<int>(B->voff[p]);
            for (int i = 0; i <= P.n; ++i) ip[r0 + i] = v0 + P.h_ip[i];
            std::vector<int> b = spmv_row_blocks(P.n, P.h_ip.data(), SPMV_THREADS, SPMV_CAP64);
            for (size_t k = 1; k < b.size(); ++k) { blocks.push_back(r0 + b[k]); bpid.push_back(p); }
            for (int64_t lo = 0; lo < P.n; lo += BCHUNK)
            {
                clo.push_back(r0 + std::min<int64_t>(lo + BCHUNK, P.n));
                cpid.push_back(p);
            }
            cptr.push_back(static_cast<int>(cpid.size()));
        }
        B->nblocks = static_cast<int>(bpid.size());
        B->nch = static_cast<int>(cpid.size());
        B->h_cptr = cptr;
        B->ip = upload(ip, s);
        B->blocks = upload(blocks, s);
        B->bpid = upload(bpid, s);
        B->clo = upload(clo, s);
        B->cpid = upload(cpid, s);
        B->cptr = upload(cptr, s);
        B->ix = DevBuf(4 * size_t(std::max<int64_t>(B->nnz, 1)));
        for (int p = 0; p < np; ++p)
        {
            const DevPattern &P = *mats[p]->pat;
            if (P.nnz)
                hipLaunchKernelGGL(k_offset_add, dim3(grid1(P.nnz)), dim3(BT), 0, s, P.nnz, P.ix.as<int>(),
                                   static_cast<int>(B->roff[p]), B->ix.as<int>() + B->voff[p]);
        }
        B->diag = DevBuf(4 * size_t(std::max<int64_t>(B->nrows, 1)));
        launch_diag_positions(static_cast<int>(B->nrows), B->ip.as<int>(), B->ix.as<int>(), B->diag.as<int>(), s);
        const size_t n = size_t(std::max<int64_t>(B->nrows, 1)), z = size_t(std::max<int64_t>(B->nnz, 1));
        B->v = DevBuf(8 * z);
        B->v32 = DevBuf(4 * z);
        B->inv32 = DevBuf(4 * n);
        B->r32 = DevBuf(4 * n); B->x32 = DevBuf(4 * n); B->p32 = DevBuf(4 * n); B->Ap32 = DevBuf(4 * n);
        B->r64 = DevBuf(8 * n); B->t64 = DevBuf(8 * n);
        B->prr = DevBuf(8 * size_t(B->nch)); B->prz = DevBuf(8 * size_t(B->nch));
        B->st = DevBuf(sizeof(BState) * np);
        B->scal = DevBuf(8 * size_t(np)); B->back = DevBuf(8 * size_t(np)); B->active = DevBuf(4 * size_t(np));
        B->cnt = DevBuf(sizeof(int));
        batch_copy_values(B.get(), s);
        *out = B.release();
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

int femd_batch_solve_refined_dev(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer,
                                 int use_graph, int *iterations, double *residuals, femd_solve_info *info, femd_stream st)
{
    return api([&] {
        if (!B) fail(FEMD_EINVAL, "null batch");
        batch_solve(B, b, x, tol, inner_tol, max_outer, use_graph, iterations, residuals, info, S(st));
    });
}

int femd_batch_solve_refined(femd_batch B, const double *b, double *x, double tol, double inner_tol, int max_outer,
                             int use_graph, int *iterations, double *residuals, femd_solve_info *info)
{
    return api([&] {
        if (!B) fail(FEMD_EINVAL, "null batch");
        hipStream_t s = S(nullptr);
        DevBuf bd = upload(b, size_t(B->nrows), s), xd = upload(x, size_t(B->nrows), s);
        batch_solve(B, bd.as<double>(), xd.as<double>(), tol, inner_tol, max_outer, use_graph, iterations, residuals, info, s);
        if (B->nrows) HIPC(hipMemcpyAsync(x, xd.p, 8 * size_t(B->nrows), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

void femd_batch_free(femd_batch B)
{
    if (B && g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
    delete B;
}

} // extern "C"
