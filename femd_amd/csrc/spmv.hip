// spmv.hip -- SpMV kernel variants and the autotuner that picks one per matrix and precision.
//
// Two families:
//  * CSR-stream <T, THREADS, CAP>: a block owns a run of whole rows whose entries fit CAP; their
//    products v[p] * x[ix[p]] are formed with coalesced loads and staged in LDS, then one thread
//    per row adds them left to right.  This is CSRMatrix::apply's order (sparse/csr.hpp:203-215),
//    so these are the only variants the exact FP64 path may use.
//  * sub-wavefront <T, LPR>: LPR lanes per row stride through the row with fused multiply-adds
//    and combine by cross-lane shuffles (no LDS, no barrier); faster on short rows, not in the
//    reference's order, so offered to the FP32 inner solves only.
// Both can also leave per-block partial sums of x_r * y_r (the p^T A p of CG) in `parts`.
//
// The autotuner times every admissible variant on the matrix itself (HIP events, a few
// repetitions each) the first time a pattern is used in a precision, and keeps the fastest.
#include "internal.hpp"

#include <cfloat>

namespace femd_amd {
namespace {

__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ inline T block_reduce(T v, T *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

template <typename T, int THREADS, int CAP, bool DOT>
__global__ __launch_bounds__(THREADS) void k_stream(const int *__restrict__ blocks, const int *__restrict__ ip,
                                                    const int *__restrict__ ix, const T *__restrict__ v,
                                                    const T *__restrict__ x, T *__restrict__ y, double *__restrict__ parts,
                                                    const int *stop)
{
    if (stop && *stop) return;
    __shared__ T prod[CAP];
    __shared__ double red[DOT ? THREADS : 1];
    const int r0 = blocks[blockIdx.x], r1 = blocks[blockIdx.x + 1];
    const int p0 = ip[r0], len = ip[r1] - p0;
    double part = 0.0;
    if (len > CAP)
    {
        // a single row longer than the staging buffer: one thread, from global memory
        if (threadIdx.x == 0)
        {
            T s = 0;
            for (int p = p0; p < p0 + len; ++p) s += v[p] * x[ix[p]];
            y[r0] = s;
            part = static_cast<double>(s) * static_cast<double>(x[r0]);
        }
    }
    else
    {
        for (int k = threadIdx.x; k < len; k += THREADS)
            prod[k] = __builtin_nontemporal_load(v + p0 + k) * x[__builtin_nontemporal_load(ix + p0 + k)];
        __syncthreads();
        const int r = r0 + threadIdx.x;
        if (r < r1)
        {
            T s = 0;
            for (int k = ip[r] - p0, e = ip[r + 1] - p0; k < e; ++k) s += prod[k];
            y[r] = s;
            if (DOT) part = static_cast<double>(s) * static_cast<double>(x[r]);
        }
    }
    if (DOT)
    {
        const double t = block_reduce(part, red);
        if (threadIdx.x == 0) parts[blockIdx.x] = t;
    }
}

template <typename T, int LPR, bool DOT>
__global__ __launch_bounds__(256) void k_subwave(int n, const int *__restrict__ ip, const int *__restrict__ ix,
                                                 const T *__restrict__ v, const T *__restrict__ x, T *__restrict__ y,
                                                 double *__restrict__ parts, const int *stop)
{
    if (stop && *stop) return;
    __shared__ double red[DOT ? 256 : 1];
    const int lane = threadIdx.x % LPR;
    const int r = blockIdx.x * (256 / LPR) + threadIdx.x / LPR;
    T s = 0;
    if (r < n)
    {
        const int e = ip[r + 1];
        for (int p = ip[r] + lane; p < e; p += LPR)
            s = fma_t(__builtin_nontemporal_load(v + p), x[__builtin_nontemporal_load(ix + p)], s);
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
    double part = 0.0;
    if (r < n && lane == 0)
    {
        y[r] = s;
        if (DOT) part = static_cast<double>(s) * static_cast<double>(x[r]);
    }
    if (DOT)
    {
        const double t = block_reduce(part, red);
        if (threadIdx.x == 0) parts[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------------- dispatch
struct Variant {
    int kind;       // 0 stream, 1 sub-wavefront
    int threads, cap, lpr;
    const char *name;
};

constexpr Variant VARIANTS[] = {
    {0, 256, 4096, 0, "stream256x4096"}, {0, 256, 2048, 0, "stream256x2048"}, {0, 512, 4096, 0, "stream512x4096"},
    {0, 128, 1024, 0, "stream128x1024"}, {1, 256, 0, 2, "subwave2"},          {1, 256, 0, 4, "subwave4"},
    {1, 256, 0, 8, "subwave8"},          {1, 256, 0, 16, "subwave16"},
};
constexpr int NVARIANTS = sizeof(VARIANTS) / sizeof(VARIANTS[0]);

template <typename T, bool DOT>
void launch_variant(const Variant &V, int nblocks, const int *blocks, int n, const int *ip, const int *ix, const T *v,
                    const T *x, T *y, double *parts, const int *stop, hipStream_t s)
{
    if (n <= 0) return;
    if (V.kind == 0)
    {
        if (V.threads == 256 && V.cap == 4096)
            hipLaunchKernelGGL((k_stream<T, 256, 4096, DOT>), dim3(nblocks), dim3(256), 0, s, blocks, ip, ix, v, x, y, parts, stop);
        else if (V.threads == 256 && V.cap == 2048)
            hipLaunchKernelGGL((k_stream<T, 256, 2048, DOT>), dim3(nblocks), dim3(256), 0, s, blocks, ip, ix, v, x, y, parts, stop);
        else if (V.threads == 512)
            hipLaunchKernelGGL((k_stream<T, 512, 4096, DOT>), dim3(nblocks), dim3(512), 0, s, blocks, ip, ix, v, x, y, parts, stop);
        else
            hipLaunchKernelGGL((k_stream<T, 128, 1024, DOT>), dim3(nblocks), dim3(128), 0, s, blocks, ip, ix, v, x, y, parts, stop);
        return;
    }
    const int g = (n + 256 / V.lpr - 1) / (256 / V.lpr);
    switch (V.lpr)
    {
    case 2: hipLaunchKernelGGL((k_subwave<T, 2, DOT>), dim3(g), dim3(256), 0, s, n, ip, ix, v, x, y, parts, stop); break;
    case 4: hipLaunchKernelGGL((k_subwave<T, 4, DOT>), dim3(g), dim3(256), 0, s, n, ip, ix, v, x, y, parts, stop); break;
    case 8: hipLaunchKernelGGL((k_subwave<T, 8, DOT>), dim3(g), dim3(256), 0, s, n, ip, ix, v, x, y, parts, stop); break;
    default: hipLaunchKernelGGL((k_subwave<T, 16, DOT>), dim3(g), dim3(256), 0, s, n, ip, ix, v, x, y, parts, stop); break;
    }
}

} // namespace

namespace {

SpmvPlan make_plan(const DevPattern &P, int vi, hipStream_t s)
{
    SpmvPlan p;
    p.variant = vi;
    const Variant &V = VARIANTS[vi];
    if (V.kind == 0)
    {
        std::vector<int> b = spmv_row_blocks(P.n, P.h_ip.data(), V.threads, V.cap);
        p.nblocks = static_cast<int>(b.size()) - 1;
        p.blocks = upload(b, s);
    }
    else p.nblocks = (P.n + 256 / V.lpr - 1) / (256 / V.lpr);
    return p;
}

template <typename T>
void run_plan(const SpmvPlan &p, const DevPattern &P, const T *v, const T *x, T *y, double *parts, const int *stop, hipStream_t s)
{
    const Variant &V = VARIANTS[p.variant];
    if (parts)
        launch_variant<T, true>(V, p.nblocks, p.blocks.as<int>(), P.n, P.ip.as<int>(), P.ix.as<int>(), v, x, y, parts, stop, s);
    else
        launch_variant<T, false>(V, p.nblocks, p.blocks.as<int>(), P.n, P.ip.as<int>(), P.ix.as<int>(), v, x, y, nullptr, stop, s);
}

// Time every admissible variant on the pattern (values v) and keep the fastest.
template <typename T>
SpmvPlan tune(const DevPattern &P, const T *v, bool exact, hipStream_t s)
{
    SpmvPlan best;
    best.ms = DBL_MAX;
    if (P.n == 0) return make_plan(P, 0, s);
    const int64_t xw = (sizeof(T) * int64_t(std::max(P.m, 1)) + 7) / 8;     // x as doubles: zero bytes are zero floats too
    DevBuf x(8 * size_t(xw)), y(sizeof(T) * std::max(P.n, 1));
    launch_fill(xw, 0.0, x.as<double>(), s);
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    for (int vi = 0; vi < NVARIANTS; ++vi)
    {
        if (exact && VARIANTS[vi].kind != 0) continue;
        SpmvPlan p = make_plan(P, vi, s);
        for (int k = 0; k < 2; ++k) run_plan<T>(p, P, v, x.as<T>(), y.as<T>(), nullptr, nullptr, s);
        const int reps = P.nnz > (int64_t(1) << 22) ? 5 : 20;
        HIPC(hipEventRecord(e0, s));
        for (int k = 0; k < reps; ++k) run_plan<T>(p, P, v, x.as<T>(), y.as<T>(), nullptr, nullptr, s);
        HIPC(hipEventRecord(e1, s));
        HIPC(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIPC(hipEventElapsedTime(&ms, e0, e1));
        p.ms = ms / reps;
        if (p.ms < best.ms) best = std::move(p);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

} // namespace

const SpmvPlan &spmv_plan(DevPattern &P, int precision, bool exact, const void *values, hipStream_t s)
{
    std::unique_ptr<SpmvPlan> &slot = precision == FEMD_FP32 ? (exact ? P.plan32x : P.plan32) : P.plan64;
    if (!slot)
    {
        slot = std::make_unique<SpmvPlan>(precision == FEMD_FP32 ? tune<float>(P, static_cast<const float *>(values), exact, s)
                                                                 : tune<double>(P, static_cast<const double *>(values), true, s));
    }
    return *slot;
}

void spmv_run64(const SpmvPlan &p, const DevPattern &P, const double *v, const double *x, double *y, double *parts,
                const int *stop, hipStream_t s)
{
    run_plan<double>(p, P, v, x, y, parts, stop, s);
}

void spmv_run32(const SpmvPlan &p, const DevPattern &P, const float *v, const float *x, float *y, double *parts,
                const int *stop, hipStream_t s)
{
    run_plan<float>(p, P, v, x, y, parts, stop, s);
}

const char *spmv_variant_name(const SpmvPlan &p) { return VARIANTS[p.variant].name; }

} // namespace femd_amd

// --------------------------------------------------------------------------------- C ABI
extern "C" int femd_spmv_tuning(femd_csr A, int precision, char *name, int name_len, double *ms)
{
    using namespace femd_amd;
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        const SpmvPlan &p = precision == FEMD_FP32 ? spmv_plan(*a->pat, FEMD_FP32, false, values32(a, s), s)
                                                   : spmv_plan(*a->pat, FEMD_FP64, true, a->v.p, s);
        if (name && name_len > 0)
        {
            std::strncpy(name, spmv_variant_name(p), size_t(name_len) - 1);
            name[name_len - 1] = '\0';
        }
        if (ms) *ms = p.ms;
    });
}
