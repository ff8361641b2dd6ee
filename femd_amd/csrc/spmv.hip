// spmv.hip -- the CSR SpMV tile, its launch shapes, the autotuner that picks one per matrix and
// precision, and the product fused into the FP32 inner CG.
//
// The tile: a run of whole rows whose entries fit the LDS buffer (CAP entries, at most THREADS
// rows), one workgroup each.  Finite-element rows of one tile share most of their columns (a
// P3 tile of 2048 entries touches about 480 distinct ones), so the tile's distinct columns are
// listed ahead of time (sorted, per pattern) and every entry carries a 16-bit position in that
// list instead of its 32-bit column.  The workgroup
//   1. loads the list and gathers x once per distinct column into LDS (xs), while every lane
//      loads groups of four consecutive entries -- values and local positions with vector loads;
//   2. forms the products v[p] * xs[lidx[p]] into LDS;
//   3. adds each row's products left to right from zero, one lane per row: the order of
//      CSRMatrix::apply (sparse/csr.hpp:203-215), so the FP64 product is bit-identical to the
//      reference's.
// Per entry the matrix stream is 8 + 2 (FP64) or 4 + 2 (FP32) bytes, and the scattered gathers
// shrink to one per distinct column: the per-lane gather, not HBM, limits a plain CSR tile here.
//
// Launch shapes (THREADS, CAP) are the tuning space; the autotuner times each on the matrix
// itself with HIP events the first time a pattern is used in a precision and keeps the fastest.
//
// The product of the FP32 CG (k_scg_spmv) is the same tile on r, writing w = A r and leaving the
// partials of r.r and w.r (and, after a residual replacement, w.p and p.s) per workgroup
// (cg32.hip).
#include "cg32.hpp"

#include <hip/hip_ext.h>

#include <cfloat>

namespace femd_amd {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Vals4;
template <> struct Vals4<float> {
    f32x4 a;
    __device__ void load(const float *v, int p) { a = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(v + p)); }
    __device__ float get(int j) const { return a[j]; }
    __device__ void set(int j, float t) { a[j] = t; }
};
template <> struct Vals4<double> {
    f64x2 a, b;
    __device__ void load(const double *v, int p)
    {
        a = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(v + p));
        b = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(v + p + 2));
    }
    __device__ double get(int j) const { return j < 2 ? a[j] : b[j - 2]; }
    __device__ void set(int j, double t) { if (j < 2) a[j] = t; else b[j - 2] = t; }
};

// One tile (workgroup blockIdx.x).  Op supplies gather(j) (the x value at global column j),
// own(r) (what the row epilogue needs of row r) and store(r, s, own) (writes row r's result and
// returns its share of the workgroup's partials, of type Op::Part).  Returns this thread's share.
// `stop` (may be NULL) is read together with the tile's bounds: a set flag ends the workgroup
// before it loads the matrix, and its latency hides behind theirs.  *stopped tells the caller.
template <typename T, int THREADS, int CAP, class Op>
__device__ inline typename Op::Part stream_tile(const int2 *__restrict__ blk, const int *__restrict__ uptr, const int *__restrict__ ucol,
                                     const unsigned short *__restrict__ lidx, const int *__restrict__ ip,
                                     const int *__restrict__ ix, const T *__restrict__ v, int64_t nnz, const Op &op, T *xs,
                                     T *prod, const int *stop, bool *stopped)
{
    constexpr int NGRP = (CAP / 4 + 1 + THREADS - 1) / THREADS;
    constexpr int NU = (CAP + THREADS - 1) / THREADS;
    const int b = blockIdx.x;
    const int2 b0 = blk[b], b1 = blk[b + 1];
    const int halt = stop ? *stop : 0;
    *stopped = halt != 0;
    if (halt) return typename Op::Part{};
    const int r0 = b0.x, r1 = b1.x, p0 = b0.y, len = b1.y - p0;
    const int r = r0 + static_cast<int>(threadIdx.x);
    int rb = 0, re = 0;
    typename Op::Own own{};
    if (r < r1)
    {
        rb = ip[r] - p0;
        re = ip[r + 1] - p0;
        own = op.own(r);
    }
    if (len > CAP)
    {
        // one row longer than the buffer: one lane, from global memory, in order
        typename Op::Part part{};
        if (threadIdx.x == 0)
        {
            T s = 0;
            for (int p = p0; p < p0 + len; ++p)
            {
                const T t = v[p] * op.gather(ix[p]);
                s += t;
            }
            part = op.store(r0, s, own);
        }
        return part;
    }
    // 1. distinct columns and entry loads
    const int u0 = uptr[b], nu = uptr[b + 1] - u0;
    int uc[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i)
    {
        const int u = static_cast<int>(threadIdx.x) + i * THREADS;
        uc[i] = u < nu ? __builtin_nontemporal_load(ucol + u0 + u) : 0;
    }
    const int a0 = p0 & ~3, e = p0 + len;
    Vals4<T> vv[NGRP];
    u16x4 cc[NGRP];
#pragma unroll
    for (int g = 0; g < NGRP; ++g)
    {
        const int a = a0 + 4 * (static_cast<int>(threadIdx.x) + g * THREADS);
        if (a < e)
        {
            if (int64_t(a) + 4 <= nnz)
            {
                vv[g].load(v, a);
                cc[g] = __builtin_nontemporal_load(reinterpret_cast<const u16x4 *>(lidx + a));
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    const bool in = int64_t(a) + j < nnz;
                    vv[g].set(j, in ? v[a + j] : T(0));
                    cc[g][j] = in ? lidx[a + j] : static_cast<unsigned short>(0);
                }
            }
        }
        else cc[g] = u16x4{0, 0, 0, 0};
    }
    T xg[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i)
    {
        const int u = static_cast<int>(threadIdx.x) + i * THREADS;
        xg[i] = u < nu ? op.gather(uc[i]) : T(0);
    }
#pragma unroll
    for (int i = 0; i < NU; ++i)
    {
        const int u = static_cast<int>(threadIdx.x) + i * THREADS;
        if (u < nu) xs[u] = xg[i];
    }
    __syncthreads();
    // 2. products
#pragma unroll
    for (int g = 0; g < NGRP; ++g)
    {
        const int a = a0 + 4 * (static_cast<int>(threadIdx.x) + g * THREADS);
        if (a < e)
        {
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int k = a + j - p0;
                if (k >= 0 && k < len) prod[k] = vv[g].get(j) * xs[cc[g][j]];
            }
        }
    }
    __syncthreads();
    // 3. row sums, left to right
    typename Op::Part part{};
    if (r < r1)
    {
        T s = 0;
        for (int k = rb; k < re; ++k) s += prod[k];
        part = op.store(r, s, own);
    }
    return part;
}

// Sum of v over the workgroup in a fixed order (wave shuffles, then the waves in order); valid in
// thread 0.
template <int THREADS>
__device__ inline double wg_sum(double v, double *red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < THREADS / 64; ++w) t += red[w];
    return t;
}

template <typename T> struct PlainOp {
    using Own = int;
    using Part = int;
    const T *x;
    T *y;
    __device__ T gather(int j) const { return x[j]; }
    __device__ Own own(int) const { return 0; }
    __device__ Part store(int r, T s, Own) const
    {
        y[r] = s;
        return 0;
    }
};

template <typename T, int THREADS, int CAP>
__global__ __launch_bounds__(THREADS) void k_spmv_tile(const int2 *__restrict__ blk, const int *__restrict__ uptr,
                                                       const int *__restrict__ ucol, const unsigned short *__restrict__ lidx,
                                                       const int *__restrict__ ip, const int *__restrict__ ix,
                                                       const T *__restrict__ v, const T *__restrict__ x, T *__restrict__ y,
                                                       int64_t nnz, const int *stop)
{
    __shared__ T xs[CAP];
    __shared__ T prod[CAP];
    bool stopped = false;
    (void)stream_tile<T, THREADS, CAP>(blk, uptr, ucol, lidx, ip, ix, v, nnz, PlainOp<T>{x, y}, xs, prod, stop, &stopped);
}

// ------------------------------------------------------------------ product of the FP32 CG
// w = A r, and over the workgroup's own rows the partials of r.r and w.r; on the first
// iteration of a segment (after the residual was replaced) also w.p and p.s, from which the
// scalar step forms p'.A p' exactly.
struct Dot4 {
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    __device__ Dot4 &operator+=(const Dot4 &o)
    {
        a += o.a;
        b += o.b;
        c += o.c;
        d += o.d;
        return *this;
    }
};
struct CGOwn {
    float r = 0.0f, p = 0.0f, s = 0.0f;
};
struct CGOp {
    using Own = CGOwn;
    using Part = Dot4;
    const float *r;
    const float *p, *sv;      // NULL unless the exact p'.A p' is wanted
    float *w;
    __device__ float gather(int j) const { return r[j]; }
    __device__ Own own(int i) const
    {
        CGOwn o;
        o.r = r[i];
        if (p)
        {
            o.p = p[i];
            o.s = sv[i];
        }
        return o;
    }
    __device__ Part store(int i, float s, Own o) const
    {
        w[i] = s;
        Dot4 d;
        d.a = static_cast<double>(o.r) * static_cast<double>(o.r);
        d.b = static_cast<double>(s) * static_cast<double>(o.r);
        d.c = static_cast<double>(s) * static_cast<double>(o.p);
        d.d = static_cast<double>(o.p) * static_cast<double>(o.s);
        return d;
    }
};

template <int THREADS, int CAP>
__global__ __launch_bounds__(THREADS) void k_scg_spmv(const int2 *__restrict__ blk, const int *__restrict__ uptr,
                                                      const int *__restrict__ ucol, const unsigned short *__restrict__ lidx,
                                                      const int *__restrict__ ip, const int *__restrict__ ix,
                                                      const float *__restrict__ v, const float *__restrict__ r,
                                                      const float *__restrict__ p, const float *__restrict__ sv,
                                                      float *__restrict__ w, int64_t nnz, double *__restrict__ parts,
                                                      const SCGState *st, int k)
{
    __shared__ float xs[CAP];
    __shared__ float prod[CAP];
    __shared__ double red[4 * (THREADS / 64)];
    const bool exact = k > 0 && st->fresh;
    bool stopped = false;
    Dot4 d = stream_tile<float, THREADS, CAP>(blk, uptr, ucol, lidx, ip, ix, v, nnz,
                                              CGOp{r, exact ? p : nullptr, exact ? sv : nullptr, w}, xs, prod, &st->stop,
                                              &stopped);
    if (stopped) return;
    // the four sums in one pass: wave shuffles, then the waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        d.a += __shfl_xor(d.a, off, 64);
        d.b += __shfl_xor(d.b, off, 64);
        d.c += __shfl_xor(d.c, off, 64);
        d.d += __shfl_xor(d.d, off, 64);
    }
    if ((threadIdx.x & 63) == 0)
    {
        const int wv = threadIdx.x >> 6;
        red[4 * wv] = d.a;
        red[4 * wv + 1] = d.b;
        red[4 * wv + 2] = d.c;
        red[4 * wv + 3] = d.d;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (int wv = 0; wv < THREADS / 64; ++wv)
            for (int j = 0; j < 4; ++j) t[j] += red[4 * wv + j];
        double2 *out = reinterpret_cast<double2 *>(parts) + 2 * blockIdx.x;
        out[0] = make_double2(t[0], t[1]);
        out[1] = make_double2(t[2], t[3]);
    }
}

// ---------------------------------------------------------------------------------- dispatch
struct Shape {
    int threads, cap;
    const char *name;
};
constexpr Shape SHAPES[] = {
    {256, 1024, "tile256x1024"}, {256, 2048, "tile256x2048"}, {512, 2048, "tile512x2048"}, {512, 4096, "tile512x4096"},
};
constexpr int NSHAPES = sizeof(SHAPES) / sizeof(SHAPES[0]);

// LDS (values and products) at most 64 KiB per workgroup.
template <typename T> bool admissible(const Shape &S) { return 2 * S.cap * sizeof(T) <= 65536; }

#define FEMD_SHAPES(X) X(256, 1024) X(256, 2048) X(512, 2048) X(512, 4096)

template <typename T>
void launch_shape(const SpmvPlan &p, const Shape &S, const int *ip, const int *ix, const T *v, const T *x, T *y, int64_t nnz,
                  const int *stop, hipStream_t s)
{
    const int2 *blk = p.blocks.as<int2>();
    const int *uptr = p.uptr.as<int>(), *ucol = p.ucol.as<int>();
    const unsigned short *lidx = p.lidx.as<unsigned short>();
#define FEMD_CASE(TH, CP)                                                                                          \
    if (S.threads == TH && S.cap == CP)                                                                            \
    {                                                                                                              \
        if constexpr (2 * CP * sizeof(T) <= 65536)                                                                 \
            hipLaunchKernelGGL((k_spmv_tile<T, TH, CP>), dim3(p.ntiles), dim3(TH), 0, s, blk, uptr, ucol, lidx, ip, ix, v, \
                               x, y, nnz, stop);                                                                   \
        return;                                                                                                    \
    }
    FEMD_SHAPES(FEMD_CASE)
#undef FEMD_CASE
}

SpmvPlan make_plan(const DevPattern &P, int si, hipStream_t s)
{
    SpmvPlan p;
    p.variant = si;
    const Shape &S = SHAPES[si];
    const std::vector<int> rows = spmv_row_blocks(P.n, P.h_ip.data(), S.threads, S.cap);
    p.ntiles = static_cast<int>(rows.size()) - 1;
    std::vector<int> b2(2 * rows.size());
    for (size_t k = 0; k < rows.size(); ++k)
    {
        b2[2 * k] = rows[k];
        b2[2 * k + 1] = P.h_ip[static_cast<size_t>(rows[k])];
    }
    p.blocks = upload(b2, s);
    std::vector<int> uptr, ucol;
    std::vector<uint16_t> lidx;
    spmv_tile_columns(P.h_ip.data(), P.h_ix.data(), rows, S.cap, uptr, ucol, lidx);
    p.uptr = upload(uptr, s);
    if (ucol.empty()) ucol.push_back(0);
    p.ucol = upload(ucol, s);
    lidx.resize(lidx.size() + 8, 0);                   // room for the last group's vector load
    p.lidx = upload(lidx, s);
    HIPC(hipStreamSynchronize(s));                     // the host vectors go out of scope
    return p;
}

template <typename T>
void run_plan(const SpmvPlan &p, const DevPattern &P, const T *v, const T *x, T *y, const int *stop, hipStream_t s)
{
    if (P.n <= 0 || p.ntiles <= 0) return;
    launch_shape<T>(p, SHAPES[p.variant], P.ip.as<int>(), P.ix.as<int>(), v, x, y, P.nnz, stop, s);
}

// Time every admissible shape on the pattern (values v, x = 0) and keep the fastest.
template <typename T>
SpmvPlan tune(const DevPattern &P, const T *v, hipStream_t s)
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
    for (int si = 0; si < NSHAPES; ++si)
    {
        if (!admissible<T>(SHAPES[si])) continue;
        SpmvPlan p = make_plan(P, si, s);
        for (int k = 0; k < 2; ++k) run_plan<T>(p, P, v, x.as<T>(), y.as<T>(), nullptr, s);
        const int reps = P.nnz > (int64_t(1) << 22) ? 5 : 20;
        HIPC(hipEventRecord(e0, s));
        for (int k = 0; k < reps; ++k) run_plan<T>(p, P, v, x.as<T>(), y.as<T>(), nullptr, s);
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

const SpmvPlan &spmv_plan(DevPattern &P, int precision, const void *values, hipStream_t s)
{
    std::unique_ptr<SpmvPlan> &slot = precision == FEMD_FP32 ? P.plan32 : P.plan64;
    if (!slot)
    {
        slot = std::make_unique<SpmvPlan>(precision == FEMD_FP32 ? tune<float>(P, static_cast<const float *>(values), s)
                                                                 : tune<double>(P, static_cast<const double *>(values), s));
    }
    return *slot;
}

void spmv_run64(const SpmvPlan &p, const DevPattern &P, const double *v, const double *x, double *y, const int *stop,
                hipStream_t s)
{
    run_plan<double>(p, P, v, x, y, stop, s);
}

void spmv_run32(const SpmvPlan &p, const DevPattern &P, const float *v, const float *x, float *y, const int *stop,
                hipStream_t s)
{
    run_plan<float>(p, P, v, x, y, stop, s);
}

int scg_spmv(const SpmvPlan &p, const DevPattern &P, const float *v, const float *r, const float *pv, const float *sv, float *w,
             double *parts, const SCGState *st, int k, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    const Shape &S = SHAPES[p.variant];
    const int2 *blk = p.blocks.as<int2>();
    const int *uptr = p.uptr.as<int>(), *ucol = p.ucol.as<int>();
    const unsigned short *lidx = p.lidx.as<unsigned short>();
    if (P.n <= 0 || p.ntiles <= 0) return 0;
#define FEMD_CASE(TH, CP)                                                                                        \
    if (S.threads == TH && S.cap == CP)                                                                          \
    {                                                                                                            \
        if (e0)                                                                                                  \
            hipExtLaunchKernelGGL((k_scg_spmv<TH, CP>), dim3(p.ntiles), dim3(TH), 0, s, e0, e1, 0, blk, uptr, ucol, lidx, \
                                  P.ip.as<int>(), P.ix.as<int>(), v, r, pv, sv, w, P.nnz, parts, st, k);         \
        else                                                                                                     \
            hipLaunchKernelGGL((k_scg_spmv<TH, CP>), dim3(p.ntiles), dim3(TH), 0, s, blk, uptr, ucol, lidx,      \
                               P.ip.as<int>(), P.ix.as<int>(), v, r, pv, sv, w, P.nnz, parts, st, k);            \
        return p.ntiles;                                                                                         \
    }
    FEMD_SHAPES(FEMD_CASE)
#undef FEMD_CASE
    return 0;
}

const char *spmv_variant_name(const SpmvPlan &p) { return SHAPES[p.variant].name; }

const int NSHAPES_PUBLIC = NSHAPES;
const char *shape_name(int i) { return SHAPES[i].name; }
int shape_index(const char *name, int precision)
{
    for (int si = 0; si < NSHAPES; ++si)
        if (std::strcmp(SHAPES[si].name, name) == 0)
            return (precision == FEMD_FP32 ? admissible<float>(SHAPES[si]) : admissible<double>(SHAPES[si])) ? si : -1;
    return -1;
}
SpmvPlan shape_plan(const DevPattern &P, int si, int, hipStream_t s) { return make_plan(P, si, s); }

} // namespace femd_amd

// --------------------------------------------------------------------------------- C ABI
extern "C" int femd_spmv_tuning(femd_csr A, int precision, char *name, int name_len, double *ms)
{
    using namespace femd_amd;
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        const SpmvPlan &p = precision == FEMD_FP32 ? spmv_plan(*a->pat, FEMD_FP32, values32(a, s), s)
                                                   : spmv_plan(*a->pat, FEMD_FP64, a->v.p, s);
        if (name && name_len > 0)
        {
            std::strncpy(name, spmv_variant_name(p), size_t(name_len) - 1);
            name[name_len - 1] = '\0';
        }
        if (ms) *ms = p.ms;
    });
}

extern "C" int femd_spmv_shape_count(void) { return femd_amd::NSHAPES_PUBLIC; }

extern "C" int femd_spmv_shape_name(int index, char *name, int name_len)
{
    using namespace femd_amd;
    return api([&] {
        if (index < 0 || index >= NSHAPES_PUBLIC) fail(FEMD_EINVAL, "SpMV shape index out of range");
        if (!name || name_len <= 0) fail(FEMD_EINVAL, "SpMV shape name: no buffer");
        std::strncpy(name, shape_name(index), size_t(name_len) - 1);
        name[name_len - 1] = '\0';
    });
}

extern "C" int femd_spmv_select(femd_csr A, int precision, const char *name)
{
    using namespace femd_amd;
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        std::unique_ptr<SpmvPlan> &slot = precision == FEMD_FP32 ? a->pat->plan32 : a->pat->plan64;
        if (!name || !*name) { slot.reset(); return; }
        const int si = shape_index(name, precision);
        if (si < 0) fail(FEMD_EINVAL, std::string("SpMV shape '") + name + "' is unknown or not available in this precision");
        slot = std::make_unique<SpmvPlan>(shape_plan(*a->pat, si, precision, s));
    });
}
