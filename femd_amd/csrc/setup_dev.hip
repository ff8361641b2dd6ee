// setup_dev.hip -- the inner CG's set-up built on the device: the first-met order of a square
// pattern's unknowns (with P A P^T's pattern and the entry map) and the per-tile column lists of an
// SpMV plan.  Both give, entry for entry, what the host builders in host.cpp give
// (first_met_order, spmv_tile_columns); the host ones remain for femd_device_setup(0).
//
// First-met order, per segment: the columns in the order a row-by-row scan first meets them, then
// the columns the scan never meets, ascending.  On the device: every column's first position is the
// smallest entry index holding it (an atomic minimum over the entries); sorting the columns by
// (first position, or the segment's end and the column for the unmet ones) gives the order, since
// a segment's entries and rows both come before the next segment's.  Row k of P A P^T is row
// perm[k] with its columns renumbered and ascending: one radix sort of every entry keyed by
// (new row, new column), carrying the entry's old index (the entry map).
//
// Tile columns: one workgroup per tile sorts the tile's column indices in LDS (bitonic), keeps the
// distinct ones and finds every entry's place among them by binary search.
#include "internal.hpp"

#include <hipcub/hipcub.hpp>

#include <atomic>
#include <climits>

namespace femd_amd {

std::atomic<int> g_device_setup{1};

namespace {

constexpr int ST = 256;                // threads per workgroup of the row kernels
constexpr int SL = 16;                 // lanes per row
constexpr int TC_THREADS = 256;
constexpr int TC_CAP = 4096;           // largest tile the column kernel takes (the largest launch shape's)

int bit_length(uint64_t v)
{
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

// the segment [r0, r1) holding row i (segments: nseg + 1 ascending bounds, 0 .. n)
__device__ inline void segment_of(int i, const int64_t *seg, int nseg, int &r0, int &r1)
{
    int lo = 0, hi = nseg;               // seg[lo] <= i < seg[hi]
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (seg[mid] <= i) lo = mid; else hi = mid;
    }
    r0 = static_cast<int>(seg[lo]);
    r1 = static_cast<int>(seg[hi]);
}

__global__ void k_fill_int(int64_t n, int v, int *a)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// first[j] = the smallest entry index holding column j; *outside = 1 when a row meets a column
// outside its segment
__global__ __launch_bounds__(ST) void k_first_met(int n, const int *ip, const int *ix, const int64_t *seg, int nseg,
                                                  int *first, int *outside)
{
    const int64_t g = int64_t(blockIdx.x) * ST + threadIdx.x;
    const int i = static_cast<int>(g / SL), l = static_cast<int>(g % SL);
    if (i >= n) return;
    int r0, r1;
    segment_of(i, seg, nseg, r0, r1);
    bool bad = false;
    for (int e = ip[i] + l; e < ip[i + 1]; e += SL)
    {
        const int j = ix[e];
        if (j < r0 || j >= r1) bad = true;
        else atomicMin(&first[j], e);
    }
    if (bad) atomicOr(outside, 1);
}

// sort keys of the columns: met ones by (first + 1) << 32, unmet ones by (the segment's end
// entry) << 32 | (j + 1), so that within a segment the met come first, by first position, then the
// unmet ascending, and every segment before the next
__global__ void k_order_keys(int n, const int *ip, const int64_t *seg, int nseg, const int *first, uint64_t *key, int *val)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int f = first[j];
    uint64_t k;
    if (f != INT_MAX) k = uint64_t(uint32_t(f) + 1u) << 32;
    else
    {
        int r0, r1;
        segment_of(j, seg, nseg, r0, r1);
        k = (uint64_t(uint32_t(ip[r1])) << 32) | uint64_t(uint32_t(j) + 1u);
    }
    key[j] = k;
    val[j] = j;
}

// iperm[perm[k]] = k and the new rows' lengths
__global__ void k_inverse_lengths(int n, const int *ip, const int *perm, int *iperm, int *len)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    if (k == n) { len[n] = 0; return; }
    const int i = perm[k];
    iperm[i] = k;
    len[k] = ip[i + 1] - ip[i];
}

// every entry keyed by (new row, new column), carrying its index
__global__ __launch_bounds__(ST) void k_entry_keys(int n, const int *ip, const int *ix, const int *iperm, uint64_t *key,
                                                   int *val)
{
    const int64_t g = int64_t(blockIdx.x) * ST + threadIdx.x;
    const int i = static_cast<int>(g / SL), l = static_cast<int>(g % SL);
    if (i >= n) return;
    const uint64_t hi = uint64_t(uint32_t(iperm[i])) << 32;
    for (int e = ip[i] + l; e < ip[i + 1]; e += SL)
    {
        key[e] = hi | uint64_t(uint32_t(iperm[ix[e]]));
        val[e] = e;
    }
}

__global__ void k_low_words(int64_t nnz, const uint64_t *key, int *ix2)
{
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nnz) ix2[e] = static_cast<int>(uint32_t(key[e]));
}

// One workgroup per tile: the tile's distinct columns, ascending, into tmp[p0 ..) and their count
// into cnt[b]; every entry's position among them into lidx.  Tiles of more than cap entries (a
// single long row) keep no list (count 0, positions 0).
__global__ __launch_bounds__(TC_THREADS) void k_tile_columns(int ntiles, const int2 *blk, const int *ix, int cap, int *cnt,
                                                             int *tmp, unsigned short *lidx)
{
    __shared__ int key[TC_CAP];
    __shared__ int uniq[TC_CAP];
    __shared__ int part[TC_THREADS];
    const int b = blockIdx.x, t = threadIdx.x;
    const int p0 = blk[b].y, p1 = blk[b + 1].y, len = p1 - p0;
    if (len > cap || len <= 0)
    {
        if (t == 0) cnt[b] = 0;
        return;
    }
    int L = 1;
    while (L < len) L <<= 1;
    for (int q = t; q < L; q += TC_THREADS) key[q] = q < len ? ix[p0 + q] : INT_MAX;
    __syncthreads();
    // bitonic sort of L keys, ascending
    for (int k = 2; k <= L; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int q = t; q < L; q += TC_THREADS)
            {
                const int o = q ^ j;
                if (o > q)
                {
                    const int a = key[q], c = key[o];
                    const bool up = (q & k) == 0;
                    if ((a > c) == up) { key[q] = c; key[o] = a; }
                }
            }
            __syncthreads();
        }
    // distinct keys: each thread takes a contiguous run of the sorted keys
    const int per = (len + TC_THREADS - 1) / TC_THREADS;
    const int q0 = min(len, t * per), q1 = min(len, q0 + per);
    int c = 0;
    for (int q = q0; q < q1; ++q) c += (q == 0 || key[q] != key[q - 1]) ? 1 : 0;
    part[t] = c;
    __syncthreads();
    // exclusive scan of the runs' counts (Hillis-Steele over the workgroup)
    for (int d = 1; d < TC_THREADS; d <<= 1)
    {
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int u = part[t] - c;
    for (int q = q0; q < q1; ++q)
        if (q == 0 || key[q] != key[q - 1]) uniq[u++] = key[q];
    const int total = part[TC_THREADS - 1];
    __syncthreads();
    for (int q = t; q < total; q += TC_THREADS) tmp[p0 + q] = uniq[q];
    if (t == 0) cnt[b] = total;
    for (int q = t; q < len; q += TC_THREADS)
    {
        const int col = ix[p0 + q];
        int lo = 0, hi = total;            // uniq[lo] <= col < uniq[hi]
        while (hi - lo > 1)
        {
            const int mid = (lo + hi) >> 1;
            if (uniq[mid] <= col) lo = mid; else hi = mid;
        }
        lidx[p0 + q] = static_cast<unsigned short>(lo);
    }
}

__global__ void k_tile_compact(int ntiles, const int2 *blk, const int *uptr, const int *tmp, int *ucol)
{
    const int b = blockIdx.x;
    const int p0 = blk[b].y, u0 = uptr[b], nu = uptr[b + 1] - u0;
    for (int q = threadIdx.x; q < nu; q += blockDim.x) ucol[u0 + q] = tmp[p0 + q];
}

// ---- the pattern of a space and its gather lists ----------------------------------------------
// every element-matrix entry e = (c nl + a) nl + b keyed by (row d[a], column d[b]), or by (n, 0)
// (after every stored entry) when either dof is constrained
__global__ void k_pair_keys(int64_t E, int nl, int n, const int *__restrict__ cd, uint64_t *key, int *val, int *count)
{
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int nn = nl * nl;
    const int c = static_cast<int>(e / nn), ab = static_cast<int>(e % nn), a = ab / nl, b = ab % nl;
    const int ra = cd[int64_t(c) * nl + a], rb = cd[int64_t(c) * nl + b];
    const bool ok = ra >= 0 && rb >= 0;
    key[e] = ok ? (uint64_t(uint32_t(ra)) << 32) | uint32_t(rb) : uint64_t(uint32_t(n)) << 32;
    val[e] = static_cast<int>(e);
    if (ok) atomicAdd(count, 1);
}

// first[t] = 1 where sorted entry t starts a stored entry (t < M), 0 past M
__global__ void k_first_flags(int64_t M, const uint64_t *__restrict__ key, int *flag)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t > M) return;
    flag[t] = t < M && (t == 0 || key[t] != key[t - 1]) ? 1 : 0;
}

// the stored entries: columns, the first element entry of each, the rows' counts
__global__ void k_emit_entries(int64_t M, const uint64_t *__restrict__ key, const int *__restrict__ flag,
                               const int *__restrict__ pos, int *ix, int *mptr, int *rowcnt)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= M) return;
    if (t == M - 1) mptr[pos[M]] = static_cast<int>(M);
    if (!flag[t]) return;
    const int p = pos[t];
    ix[p] = static_cast<int>(uint32_t(key[t]));
    mptr[p] = static_cast<int>(t);
    atomicAdd(&rowcnt[key[t] >> 32], 1);
}

// every (cell, local dof) keyed by its dof, n when constrained
__global__ void k_dof_keys(int64_t L, int n, const int *__restrict__ cd, uint32_t *key, int *val, int *count, int *dofcnt)
{
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= L) return;
    const int d = cd[e];
    key[e] = d >= 0 ? uint32_t(d) : uint32_t(n);
    val[e] = static_cast<int>(e);
    if (d >= 0)
    {
        atomicAdd(count, 1);
        atomicAdd(&dofcnt[d], 1);
    }
}

template <class K, class V>
void sort_pairs(const K *kin, K *kout, const V *vin, V *vout, int64_t n, int end_bit, hipStream_t s)
{
    size_t bytes = 0;
    HIPC(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, static_cast<int>(n), 0, end_bit, s));
    DevBuf tmp(std::max<size_t>(bytes, 8));
    HIPC(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin, kout, vin, vout, static_cast<int>(n), 0, end_bit, s));
    HIPC(hipStreamSynchronize(s));         // tmp returns to the pool
}

void exclusive_sum(const int *in, int *out, int64_t n, hipStream_t s)
{
    size_t bytes = 0;
    HIPC(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, static_cast<int>(n), s));
    DevBuf tmp(std::max<size_t>(bytes, 8));
    HIPC(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, static_cast<int>(n), s));
    HIPC(hipStreamSynchronize(s));
}

int grid(int64_t n, int threads) { return static_cast<int>((n + threads - 1) / threads); }

} // namespace

bool device_setup_enabled() { return g_device_setup.load() != 0; }

bool device_first_met_order(int n, const int *ip, const int *ix, int64_t nnz, const std::vector<int64_t> &segments,
                            LocalOrder &L, hipStream_t s)
{
    if (nnz >= (int64_t(1) << 31) - 1) fail(FEMD_EINVAL, "inner CG order: more than 2^31 - 2 entries");
    const std::vector<int64_t> seg = segments.empty() ? std::vector<int64_t>{0, int64_t(n)} : segments;
    const int nseg = static_cast<int>(seg.size()) - 1;
    L.perm = DevBuf(sizeof(int) * size_t(std::max(n, 1)));
    L.emap = DevBuf(sizeof(int) * size_t(std::max<int64_t>(nnz, 1)));
    auto Q = std::make_shared<DevPattern>();
    Q->n = Q->m = n;
    Q->nnz = nnz;
    Q->segments = segments;
    Q->ip = DevBuf(sizeof(int) * (size_t(n) + 1));
    Q->ix = DevBuf(sizeof(int) * size_t(std::max<int64_t>(nnz, 1)));
    Q->h_ip.assign(size_t(n) + 1, 0);
    if (n == 0)
    {
        HIPC(hipMemcpyAsync(Q->ip.p, Q->h_ip.data(), sizeof(int), hipMemcpyHostToDevice, s));
        HIPC(hipStreamSynchronize(s));
        L.pat = std::move(Q);
        return true;
    }
    DevBuf dseg = upload(seg, s);
    DevBuf first(sizeof(int) * size_t(n)), flag(sizeof(int));
    hipLaunchKernelGGL(k_fill_int, dim3(grid(n, 256)), dim3(256), 0, s, int64_t(n), INT_MAX, first.as<int>());
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, int64_t(1), 0, flag.as<int>());
    hipLaunchKernelGGL(k_first_met, dim3(grid(int64_t(n) * SL, ST)), dim3(ST), 0, s, n, ip, ix, dseg.as<int64_t>(), nseg,
                       first.as<int>(), flag.as<int>());
    int outside = 0;
    HIPC(hipMemcpyAsync(&outside, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (outside) return false;
    {
        DevBuf k0(8 * size_t(n)), k1(8 * size_t(n)), v0(4 * size_t(n));
        hipLaunchKernelGGL(k_order_keys, dim3(grid(n, 256)), dim3(256), 0, s, n, ip, dseg.as<int64_t>(), nseg, first.as<int>(),
                           k0.as<uint64_t>(), v0.as<int>());
        sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<int>(), L.perm.as<int>(), n,
                   std::min(64, 32 + bit_length(uint64_t(nnz) + 1)), s);
    }
    DevBuf iperm(sizeof(int) * size_t(n)), len(sizeof(int) * (size_t(n) + 1));
    hipLaunchKernelGGL(k_inverse_lengths, dim3(grid(int64_t(n) + 1, 256)), dim3(256), 0, s, n, ip, L.perm.as<int>(),
                       iperm.as<int>(), len.as<int>());
    exclusive_sum(len.as<int>(), Q->ip.as<int>(), int64_t(n) + 1, s);
    if (nnz > 0)
    {
        DevBuf k0(8 * size_t(nnz)), k1(8 * size_t(nnz)), v0(4 * size_t(nnz));
        hipLaunchKernelGGL(k_entry_keys, dim3(grid(int64_t(n) * SL, ST)), dim3(ST), 0, s, n, ip, ix, iperm.as<int>(),
                           k0.as<uint64_t>(), v0.as<int>());
        sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<int>(), L.emap.as<int>(), nnz,
                   std::min(64, 32 + bit_length(uint64_t(n))), s);
        hipLaunchKernelGGL(k_low_words, dim3(grid(nnz, 256)), dim3(256), 0, s, nnz, k1.as<uint64_t>(), Q->ix.as<int>());
        HIPC(hipStreamSynchronize(s));
    }
    HIPC(hipMemcpyAsync(Q->h_ip.data(), Q->ip.p, sizeof(int) * (size_t(n) + 1), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    L.pat = std::move(Q);
    return true;
}

void device_tile_columns(const int *ix, int64_t nnz, int ntiles, const int2 *blocks, int cap, DevBuf &uptr, DevBuf &ucol,
                         DevBuf &lidx, hipStream_t s)
{
    if (cap > TC_CAP) fail(FEMD_EINVAL, "tile columns: cap above " + std::to_string(TC_CAP));
    const int64_t lwords = (2 * (nnz + 8) + 3) / 4;
    lidx = DevBuf(4 * size_t(lwords));                 // nnz + 8 positions (room for the last group's vector load)
    hipLaunchKernelGGL(k_fill_int, dim3(grid(lwords, 256)), dim3(256), 0, s, lwords, 0, lidx.as<int>());
    uptr = DevBuf(sizeof(int) * (size_t(ntiles) + 1));
    if (ntiles <= 0)
    {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, int64_t(1), 0, uptr.as<int>());
        ucol = DevBuf(sizeof(int));
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, int64_t(1), 0, ucol.as<int>());
        HIPC(hipStreamSynchronize(s));
        return;
    }
    DevBuf cnt(sizeof(int) * (size_t(ntiles) + 1)), tmp(sizeof(int) * size_t(std::max<int64_t>(nnz, 1)));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, int64_t(1), 0, cnt.as<int>() + ntiles);
    hipLaunchKernelGGL(k_tile_columns, dim3(ntiles), dim3(TC_THREADS), 0, s, ntiles, blocks, ix, cap, cnt.as<int>(),
                       tmp.as<int>(), lidx.as<unsigned short>());
    exclusive_sum(cnt.as<int>(), uptr.as<int>(), int64_t(ntiles) + 1, s);
    int total = 0;
    HIPC(hipMemcpyAsync(&total, uptr.as<int>() + ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    ucol = DevBuf(sizeof(int) * (size_t(total) + 1));  // and one valid index for a tile's clamped loads
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, int64_t(1), 0, ucol.as<int>() + total);
    hipLaunchKernelGGL(k_tile_compact, dim3(ntiles), dim3(256), 0, s, ntiles, blocks, uptr.as<int>(), tmp.as<int>(),
                       ucol.as<int>());
    HIPC(hipStreamSynchronize(s));                     // cnt, tmp return to the pool
}

void device_build_pattern(int n, int ncells, int nl, const int *cell_dof_dev, DevPatternLists &out, hipStream_t s)
{
    const int64_t E = int64_t(ncells) * nl * nl, L = int64_t(ncells) * nl;
    if (E >= (int64_t(1) << 31)) fail(FEMD_EINVAL, "pattern: element matrices exceed 2^31 entries");
    out.n = n;
    DevBuf counter(2 * sizeof(int));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, int64_t(2), 0, counter.as<int>());
    // the stored entries and, for each, its element entries in ascending order
    int M = 0;
    {
        DevBuf k0(8 * size_t(std::max<int64_t>(E, 1))), k1(8 * size_t(std::max<int64_t>(E, 1)));
        DevBuf v0(4 * size_t(std::max<int64_t>(E, 1)));
        out.mlist = DevBuf(4 * size_t(std::max<int64_t>(E, 1)));
        if (E > 0)
        {
            hipLaunchKernelGGL(k_pair_keys, dim3(grid(E, 256)), dim3(256), 0, s, E, nl, n, cell_dof_dev, k0.as<uint64_t>(),
                               v0.as<int>(), counter.as<int>());
            sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<int>(), out.mlist.as<int>(), E,
                       std::min(64, 32 + bit_length(uint64_t(n))), s);
        }
        HIPC(hipMemcpyAsync(&M, counter.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        out.ncon = M;
        DevBuf flag(4 * (size_t(M) + 1)), pos(4 * (size_t(M) + 1)), rowcnt(4 * (size_t(n) + 1));
        hipLaunchKernelGGL(k_fill_int, dim3(grid(int64_t(n) + 1, 256)), dim3(256), 0, s, int64_t(n) + 1, 0, rowcnt.as<int>());
        hipLaunchKernelGGL(k_first_flags, dim3(grid(int64_t(M) + 1, 256)), dim3(256), 0, s, int64_t(M), k1.as<uint64_t>(),
                           flag.as<int>());
        exclusive_sum(flag.as<int>(), pos.as<int>(), int64_t(M) + 1, s);
        int nnz = 0;
        HIPC(hipMemcpyAsync(&nnz, pos.as<int>() + M, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        out.nnz = nnz;
        out.ix = DevBuf(4 * size_t(std::max(nnz, 1)));
        out.mptr = DevBuf(4 * (size_t(nnz) + 1));
        if (M == 0) hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, int64_t(1), 0, out.mptr.as<int>());
        else
            hipLaunchKernelGGL(k_emit_entries, dim3(grid(M, 256)), dim3(256), 0, s, int64_t(M), k1.as<uint64_t>(),
                               flag.as<int>(), pos.as<int>(), out.ix.as<int>(), out.mptr.as<int>(), rowcnt.as<int>());
        out.ip = DevBuf(4 * (size_t(n) + 1));
        exclusive_sum(rowcnt.as<int>(), out.ip.as<int>(), int64_t(n) + 1, s);
    }
    // every dof's (cell, local dof) entries in ascending order
    {
        DevBuf k0(4 * size_t(std::max<int64_t>(L, 1))), k1(4 * size_t(std::max<int64_t>(L, 1)));
        DevBuf v0(4 * size_t(std::max<int64_t>(L, 1))), dofcnt(4 * (size_t(n) + 1));
        out.vlist = DevBuf(4 * size_t(std::max<int64_t>(L, 1)));
        hipLaunchKernelGGL(k_fill_int, dim3(grid(int64_t(n) + 1, 256)), dim3(256), 0, s, int64_t(n) + 1, 0, dofcnt.as<int>());
        if (L > 0)
        {
            hipLaunchKernelGGL(k_dof_keys, dim3(grid(L, 256)), dim3(256), 0, s, L, n, cell_dof_dev, k0.as<uint32_t>(),
                               v0.as<int>(), counter.as<int>() + 1, dofcnt.as<int>());
            sort_pairs(k0.as<uint32_t>(), k1.as<uint32_t>(), v0.as<int>(), out.vlist.as<int>(), L,
                       std::max(1, bit_length(uint64_t(n))), s);
        }
        out.vptr = DevBuf(4 * (size_t(n) + 1));
        exclusive_sum(dofcnt.as<int>(), out.vptr.as<int>(), int64_t(n) + 1, s);
        HIPC(hipStreamSynchronize(s));
    }
    out.h_ip.assign(size_t(n) + 1, 0);
    out.h_ix.assign(size_t(out.nnz), 0);
    HIPC(hipMemcpyAsync(out.h_ip.data(), out.ip.p, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost, s));
    if (out.nnz) HIPC(hipMemcpyAsync(out.h_ix.data(), out.ix.p, 4 * size_t(out.nnz), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
}

} // namespace femd_amd

extern "C" {

int femd_device_setup(int mode)
{
    if (mode < 0) return femd_amd::g_device_setup.load();
    return femd_amd::g_device_setup.exchange(mode ? 1 : 0);
}

int femd_inner_order_pattern_gpu(int n, const int *indptr, const int *indices, int nseg, const int64_t *segments, int *perm,
                                 int *indptr2, int *indices2, int *emap)
{
    using namespace femd_amd;
    return api([&] {
        if (n < 0 || nseg < 0 || !indptr || (n > 0 && !indices) || !perm || !indptr2 || (nseg > 0 && !segments))
            fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: bad arguments");
        if (indptr[0] != 0) fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: indptr[0] must be 0");
        for (int i = 0; i < n; ++i)
            if (indptr[i + 1] < indptr[i]) fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: indptr decreases");
        const int64_t nnz = n > 0 ? indptr[n] : 0;
        if (nnz > 0 && (!indices2 || !emap)) fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: null output");
        std::vector<int64_t> seg;
        if (nseg > 0)
        {
            seg.assign(segments, segments + nseg + 1);
            if (seg.front() != 0 || seg.back() != n) fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: segments must cover the rows");
            for (int k = 0; k < nseg; ++k)
                if (seg[size_t(k) + 1] < seg[size_t(k)]) fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: segments decrease");
        }
        for (int64_t p = 0; p < nnz; ++p)
            if (indices[p] < 0 || indices[p] >= n) fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: column index out of range");
        hipStream_t s = S(nullptr);
        DevBuf ip = upload(indptr, size_t(n) + 1, s), ix = upload(indices, size_t(nnz), s);
        LocalOrder L;
        if (!device_first_met_order(n, ip.as<int>(), ix.as<int>(), nnz, seg, L, s))
            fail(FEMD_EINVAL, "femd_inner_order_pattern_gpu: a segment's rows have columns outside it");
        if (n > 0) HIPC(hipMemcpyAsync(perm, L.perm.p, sizeof(int) * size_t(n), hipMemcpyDeviceToHost, s));
        std::copy(L.pat->h_ip.begin(), L.pat->h_ip.end(), indptr2);
        if (nnz > 0)
        {
            HIPC(hipMemcpyAsync(indices2, L.pat->ix.p, sizeof(int) * size_t(nnz), hipMemcpyDeviceToHost, s));
            HIPC(hipMemcpyAsync(emap, L.emap.p, sizeof(int) * size_t(nnz), hipMemcpyDeviceToHost, s));
        }
        HIPC(hipStreamSynchronize(s));
    });
}

} // extern "C"
