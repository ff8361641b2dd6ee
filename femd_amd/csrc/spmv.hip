// spmv.hip -- the CSR SpMV tile, its launch shapes, the autotuner that picks one per matrix and
// precision.
//
// The tile kernels themselves are in tile.hpp.
//
// Launch shapes (THREADS, CAP) are the tuning space; the autotuner times each on the matrix
// itself with HIP events the first time a pattern is used in a precision and keeps the fastest.
//
// The product of the inner CG (cg_engine.hip) is the same tile with the direction update fused in.
#include "tile.hpp"

#include <atomic>
#include <cfloat>

namespace femd_amd {
namespace {
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


template <typename T>
void launch_shape(const SpmvPlan &p, const Shape &S, const int *ip, const int *ix, const T *v, const T *x, T *y, int64_t nnz,
                  const int *stop, const int *mask, hipStream_t s)
{
    const int *tile_sys = mask ? p.tile_sys.as<int>() : nullptr;
    const int2 *blk = p.blocks.as<int2>();
    const int *uptr = p.uptr.as<int>(), *ucol = p.ucol.as<int>();
    const unsigned short *lidx = p.lidx.as<unsigned short>();
#define FEMD_CASE(TH, CP)                                                                                          \
    if (S.threads == TH && S.cap == CP)                                                                            \
    {                                                                                                              \
        if constexpr (2 * CP * sizeof(T) <= 65536)                                                                 \
            hipLaunchKernelGGL((k_spmv_tile<T, TH, CP>), dim3(p.ntiles), dim3(TH), 0, s, blk, uptr, ucol, lidx, ip, ix, v, \
                               x, y, nnz, stop, tile_sys, mask);                                                   \
        return;                                                                                                    \
    }
    FEMD_SHAPES(FEMD_CASE)
#undef FEMD_CASE
}

SpmvPlan make_plan(const DevPattern &P, int si, hipStream_t s)
{
    static std::atomic<uint64_t> built{0};
    SpmvPlan p;
    p.id = ++built;
    p.variant = si;
    const Shape &S = SHAPES[si];
    // tiles never straddle a segment boundary
    std::vector<int64_t> seg = P.segments;
    if (seg.empty()) seg = {0, P.n};
    std::vector<int> rows{0}, tsys;
    p.seg_tiles.assign(1, 0);
    for (size_t k = 0; k + 1 < seg.size(); ++k)
    {
        const int r0 = static_cast<int>(seg[k]), nr = static_cast<int>(seg[k + 1] - seg[k]);
        const std::vector<int> b = spmv_row_blocks(nr, P.h_ip.data() + r0, S.threads, S.cap);
        for (size_t j = 1; j < b.size(); ++j)
        {
            rows.push_back(r0 + b[j]);
            tsys.push_back(static_cast<int>(k));
        }
        p.seg_tiles.push_back(static_cast<int>(tsys.size()));
    }
    p.ntiles = static_cast<int>(rows.size()) - 1;
    if (!P.segments.empty()) p.tile_sys = upload(tsys.empty() ? std::vector<int>{0} : tsys, s);
    std::vector<int> b2(2 * rows.size());
    for (size_t k = 0; k < rows.size(); ++k)
    {
        b2[2 * k] = rows[k];
        b2[2 * k + 1] = P.h_ip[static_cast<size_t>(rows[k])];
    }
    p.blocks = upload(b2, s);
    if (device_setup_enabled())
    {
        device_tile_columns(P.ix.as<int>(), P.nnz, p.ntiles, p.blocks.as<int2>(), S.cap, p.uptr, p.ucol, p.lidx, s);
        return p;
    }
    std::vector<int> uptr, ucol;
    std::vector<uint16_t> lidx;
    spmv_tile_columns(P.h_ip.data(), P.h_ix.data(), rows, S.cap, uptr, ucol, lidx);
    p.uptr = upload(uptr, s);
    ucol.push_back(0);                                 // a valid index for a tile's clamped loads
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
    launch_shape<T>(p, SHAPES[p.variant], P.ip.as<int>(), P.ix.as<int>(), v, x, y, P.nnz, stop, nullptr, s);
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

int shape_index(const char *name, int precision);

const SpmvPlan &spmv_plan(DevPattern &P, int precision, const void *values, hipStream_t s)
{
    std::unique_ptr<SpmvPlan> &slot = precision == FEMD_FP32 ? P.plan32 : P.plan64;
    if (!slot && !P.segments.empty())
    {
        // segmented (batched) patterns use fixed shapes, so that a segment is tiled -- and its
        // dot products summed -- the same way whatever else is in the batch
        // (measured on the 256-system P2 batch: FP64 256x1024 2.36 s, 512x2048 2.36 s, 256x2048
        // 2.55 s; FP32 256x2048 4.78 s, 512x4096 5.07 s, 512x2048 6.32 s; in the inner CG's order
        // of the unknowns FP64 256x1024 1.99 s, 512x2048 2.06 s, 256x2048 2.28 s, FP32 256x2048
        // 4.0 s, 512x4096 4.27 s, 256x1024 5.47 s)
        slot = std::make_unique<SpmvPlan>(make_plan(P, shape_index(precision == FEMD_FP32 ? "tile256x2048" : "tile256x1024",
                                                                   precision), s));
    }
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

void spmv_run64_masked(const SpmvPlan &p, const DevPattern &P, const double *v, const double *x, double *y, const int *mask,
                       hipStream_t s)
{
    if (P.n <= 0 || p.ntiles <= 0) return;
    if (!p.tile_sys.p) fail(FEMD_EINVAL, "masked SpMV on a pattern without segments");
    launch_shape<double>(p, SHAPES[p.variant], P.ip.as<int>(), P.ix.as<int>(), v, x, y, P.nnz, nullptr, mask, s);
}

void plan_shape(const SpmvPlan &p, int *threads, int *cap)
{
    *threads = SHAPES[p.variant].threads;
    *cap = SHAPES[p.variant].cap;
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

// ---------------------------------------------------------------------------- value dictionaries
// Built on the device in two passes over the tiles (one workgroup each): both insert the tile's
// values, by bit pattern, into an LDS hash table (open addressing, 512 slots); the first counts
// the distinct ones, the second -- once every tile is known to have at most 256 and the offsets
// are set -- numbers the occupied slots in slot order, writes the dictionary and every entry's
// position.  Values are compared by bit pattern, so the dictionary is exact (-0.0 and 0.0 stay
// apart); the one pattern used to mark empty slots (all ones, a NaN) makes a tile unusable.
namespace femd_amd {
namespace {

constexpr int DICT_SLOTS = 512;
constexpr int DICT_THREADS = 256;
constexpr unsigned long long DICT_EMPTY = ~0ull;
constexpr int DICT_OVERFLOW = 1 << 20;
constexpr int DICT_FULL = 384;             // a tile with more distinct values stops inserting (overflow)

__device__ inline unsigned long long dict_key(float v) { return __float_as_uint(v); }
__device__ inline unsigned long long dict_key(double v) { return static_cast<unsigned long long>(__double_as_longlong(v)); }
template <typename T> __device__ inline T dict_value(unsigned long long k);
template <> __device__ inline float dict_value<float>(unsigned long long k) { return __uint_as_float(static_cast<unsigned>(k)); }
template <> __device__ inline double dict_value<double>(unsigned long long k) { return __longlong_as_double(static_cast<long long>(k)); }
__device__ inline int dict_hash(unsigned long long k) { return static_cast<int>((k * 0x9E3779B97F4A7C15ull) >> 55); }

// Insert the values of tile b into keys; false (in every thread) when the table or a key cannot hold
// them, or when the tile has more than DICT_FULL distinct values (the inserts stop there: a nearly
// full table makes every further insert probe most of it, and such a tile is beyond any dictionary).
template <typename T>
__device__ bool dict_fill(int p0, int p1, const T *__restrict__ v, unsigned long long *keys, int *bad, int *filled)
{
    for (int i = threadIdx.x; i < DICT_SLOTS; i += DICT_THREADS) keys[i] = DICT_EMPTY;
    if (threadIdx.x == 0)
    {
        *bad = 0;
        *filled = 0;
    }
    __syncthreads();
    for (int p = p0 + static_cast<int>(threadIdx.x); p < p1; p += DICT_THREADS)
    {
        if (*static_cast<volatile int *>(bad)) break;
        const unsigned long long k = dict_key(v[p]);
        if (k == DICT_EMPTY) { *bad = 1; continue; }
        int h = dict_hash(k);
        for (int probe = 0;; ++probe)
        {
            if (probe == DICT_SLOTS) { *bad = 1; break; }
            const unsigned long long prev = atomicCAS(&keys[h], DICT_EMPTY, k);
            if (prev == DICT_EMPTY)
            {
                if (atomicAdd(filled, 1) >= DICT_FULL) *bad = 1;
                break;
            }
            if (prev == k) break;
            h = (h + 1) & (DICT_SLOTS - 1);
        }
    }
    __syncthreads();
    const bool ok = *bad == 0;
    __syncthreads();
    return ok;
}

template <typename T>
__global__ __launch_bounds__(DICT_THREADS) void k_dict_count(const int2 *__restrict__ blk, const T *__restrict__ v, int *counts)
{
    __shared__ unsigned long long keys[DICT_SLOTS];
    __shared__ int bad, filled;
    const int b = blockIdx.x;
    const bool ok = dict_fill(blk[b].y, blk[b + 1].y, v, keys, &bad, &filled);
    int c = 0;
    for (int i = threadIdx.x; i < DICT_SLOTS; i += DICT_THREADS) c += keys[i] != DICT_EMPTY ? 1 : 0;
    const int total = __syncthreads_count(c >= 1) + __syncthreads_count(c >= 2);
    if (threadIdx.x == 0) counts[b] = ok ? total : DICT_OVERFLOW;
}

template <typename T>
__global__ __launch_bounds__(DICT_THREADS) void k_dict_write(const int2 *__restrict__ blk, const T *__restrict__ v,
                                                             const int *__restrict__ doff, T *dval, unsigned char *vidx)
{
    __shared__ unsigned long long keys[DICT_SLOTS];
    __shared__ int bad, filled;
    __shared__ int scan[DICT_THREADS];
    __shared__ unsigned char pos[DICT_SLOTS];
    const int b = blockIdx.x, p0 = blk[b].y, p1 = blk[b + 1].y;
    if (doff[b + 1] == doff[b]) return;        // a tile with too many values reads the value array
    if (!dict_fill(p0, p1, v, keys, &bad, &filled)) return;
    // slots 2t and 2t + 1 belong to thread t; number the occupied ones in slot order
    const int t = threadIdx.x;
    const bool o0 = keys[2 * t] != DICT_EMPTY, o1 = keys[2 * t + 1] != DICT_EMPTY;
    scan[t] = int(o0) + int(o1);
    __syncthreads();
    for (int off = 1; off < DICT_THREADS; off <<= 1)
    {
        const int add = t >= off ? scan[t - off] : 0;
        __syncthreads();
        scan[t] += add;
        __syncthreads();
    }
    const int first = scan[t] - int(o0) - int(o1), d0 = doff[b];
    if (o0)
    {
        pos[2 * t] = static_cast<unsigned char>(first);
        dval[d0 + first] = dict_value<T>(keys[2 * t]);
    }
    if (o1)
    {
        pos[2 * t + 1] = static_cast<unsigned char>(first + int(o0));
        dval[d0 + first + int(o0)] = dict_value<T>(keys[2 * t + 1]);
    }
    __syncthreads();
    for (int p = p0 + t; p < p1; p += DICT_THREADS)
    {
        const unsigned long long k = dict_key(v[p]);
        int h = dict_hash(k);
        while (keys[h] != k) h = (h + 1) & (DICT_SLOTS - 1);
        vidx[p] = pos[h];
    }
}

std::atomic<int> g_dict_on{1};

} // namespace

bool dict_enabled() { return g_dict_on.load() != 0; }
std::atomic<int> g_order_on{1};
bool inner_order_enabled() { return g_order_on.load() != 0; }

void build_dict(const SpmvPlan &p, const void *v, int precision, int64_t nnz, ValueDict &d, hipStream_t s)
{
    d.valid = true;
    d.usable = false;
    d.gen += 1;
    d.plan_id = p.id;
    d.max_distinct = 0;
    d.distinct = 0;
    d.doff.release();
    d.dval.release();
    d.vidx.release();
    if (p.ntiles <= 0 || nnz <= 0) return;
    const int nt = p.ntiles;
    const int2 *blk = p.blocks.as<int2>();
    DevBuf counts(4 * size_t(nt));
    if (precision == FEMD_FP64)
        hipLaunchKernelGGL(k_dict_count<double>, dim3(nt), dim3(DICT_THREADS), 0, s, blk, static_cast<const double *>(v),
                           counts.as<int>());
    else
        hipLaunchKernelGGL(k_dict_count<float>, dim3(nt), dim3(DICT_THREADS), 0, s, blk, static_cast<const float *>(v),
                           counts.as<int>());
    std::vector<int> h(size_t(nt) + 1, 0);
    std::vector<int2> hb(size_t(nt) + 1);
    HIPC(hipMemcpyAsync(h.data(), counts.p, 4 * size_t(nt), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(hb.data(), blk, sizeof(int2) * (size_t(nt) + 1), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    // a tile with more than 256 distinct values keeps reading the value array (an empty
    // dictionary); the format pays while such tiles hold at most a quarter of the entries
    int mx = 0;
    int64_t tot = 0, raw_entries = 0;
    d.raw_tiles = 0;
    for (int b = 0; b < nt; ++b)
    {
        mx = std::max(mx, std::min(h[size_t(b)], DICT_FULL + 1));     // DICT_FULL + 1: more than DICT_FULL
        if (h[size_t(b)] > 256)
        {
            h[size_t(b)] = 0;
            raw_entries += hb[size_t(b) + 1].y - hb[size_t(b)].y;
            ++d.raw_tiles;
        }
        tot += h[size_t(b)];
    }
    d.max_distinct = mx;
    if (4 * raw_entries > nnz || tot >= (int64_t(1) << 31)) return;
    // exclusive offsets
    int run = 0;
    for (int b = 0; b < nt; ++b)
    {
        const int c = h[size_t(b)];
        h[size_t(b)] = run;
        run += c;
    }
    h[size_t(nt)] = run;
    const size_t w = precision == FEMD_FP64 ? 8 : 4;
    d.doff = upload(h, s);
    d.dval = DevBuf(w * (size_t(run) + 1));            // + 1: a valid index for a tile's clamped load
    d.vidx = DevBuf(size_t(nnz) + 8);
    if (precision == FEMD_FP64)
        hipLaunchKernelGGL(k_dict_write<double>, dim3(nt), dim3(DICT_THREADS), 0, s, blk, static_cast<const double *>(v),
                           d.doff.as<int>(), d.dval.as<double>(), d.vidx.as<unsigned char>());
    else
        hipLaunchKernelGGL(k_dict_write<float>, dim3(nt), dim3(DICT_THREADS), 0, s, blk, static_cast<const float *>(v),
                           d.doff.as<int>(), d.dval.as<float>(), d.vidx.as<unsigned char>());
    HIPC(hipGetLastError());
    d.distinct = run;
    d.usable = true;
}

} // namespace femd_amd

// --------------------------------------------------------------------------------- C ABI
extern "C" int femd_inner_order(int mode)
{
    using namespace femd_amd;
    const int prev = g_order_on.load();
    if (mode == 0 || mode == 1) g_order_on.store(mode);
    return prev;
}

extern "C" int femd_value_dictionary(int mode)
{
    using namespace femd_amd;
    const int prev = g_dict_on.load();
    if (mode == 0 || mode == 1) g_dict_on.store(mode);
    return prev;
}

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
        const int si = name && *name ? shape_index(name, precision) : -1;
        if (name && *name && si < 0)
            fail(FEMD_EINVAL, std::string("SpMV shape '") + name + "' is unknown or not available in this precision");
        // the inner CG's reordered pattern takes the same shape (or is tuned again) at its next use
        a->pat->forced[precision == FEMD_FP32 ? 1 : 0] = si;
        if (a->pat->local) (precision == FEMD_FP32 ? a->pat->local->pat->plan32 : a->pat->local->pat->plan64).reset();
        if (si < 0) { slot.reset(); return; }
        slot = std::make_unique<SpmvPlan>(shape_plan(*a->pat, si, precision, s));
    });
}
