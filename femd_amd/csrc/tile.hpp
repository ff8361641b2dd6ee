// tile.hpp -- the CSR SpMV tile shared by the single-matrix kernels (spmv.hip) and the batched
// solver (batch.hip): device code only, instantiated in each translation unit that uses it.
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
#pragma once
#include "internal.hpp"

namespace femd_amd {
namespace {

// The launch shapes (THREADS, CAP) of the tuning space.
#define FEMD_SHAPES(X) X(256, 1024) X(256, 2048) X(512, 2048) X(512, 4096)


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

// Where a tile's matrix values come from: the value array itself (NoDict), or a per-tile value
// dictionary -- the tile's distinct values (at most 256) in dval[doff[b] .. doff[b + 1]), staged
// in LDS (dv), and every entry's 1-byte position in them (vidx).  A tile with more distinct values
// has an empty dictionary and reads the value array itself.  Every entry gets its exact value
// either way, so the formats give bit-identical products.
struct NoDict {
    static constexpr bool on = false;
    static constexpr bool raw_tiles = false;
};
// RAW: some tiles have empty dictionaries and read the value array (a branch per workgroup)
template <typename T, bool RAW = false> struct TileDict {
    static constexpr bool on = true;
    static constexpr bool raw_tiles = RAW;
    const unsigned char *vidx;
    const int *doff;
    const T *dval;
    T *dv;          // LDS, 256 entries
};

typedef unsigned int u32;

// A tile's bounds: rows [r0, r1), entries [p0, p1), distinct columns ucol[u0 .. u1).
struct TileDesc {
    int r0, p0, u0, r1, p1, u1;
};

// The work of one tile b (the whole workgroup) once its bounds are known.  Op supplies gather(j)
// (the x value at global column j), own(r) (what the row epilogue needs of row r), store(r, s, own)
// (writes row r's result and returns its share of the workgroup's partials, of type Op::Part) and
// ready() (false: the tile has nothing to do -- asked once the tile's loads are in flight, so a
// flag read ahead of them hides behind their latency; the same answer in every thread).  Returns
// this thread's share.
template <typename T, int THREADS, int CAP, class Op, class VS = NoDict>
__device__ inline typename Op::Part tile_work(const TileDesc &td, int b, const int *__restrict__ ucol,
                                              const unsigned short *__restrict__ lidx, const int *__restrict__ ip,
                                              const int *__restrict__ ix, const T *__restrict__ v, int64_t nnz, const Op &op,
                                              T *xs, T *prod, const VS &vs = VS{})
{
    constexpr int NGRP = (CAP / 4 + 1 + THREADS - 1) / THREADS;
    constexpr int NU = (CAP + THREADS - 1) / THREADS;
    static_assert(!VS::on || THREADS >= 256, "a dictionary of 256 values is staged one per lane");
    const int r0 = td.r0, r1 = td.r1, p0 = td.p0, len = td.p1 - td.p0;
    const int r = r0 + static_cast<int>(threadIdx.x);
    int d0 = 0, nd = 0;
    if constexpr (VS::on)
    {
        d0 = vs.doff[b];
        nd = vs.doff[b + 1] - d0;
    }
    // the tile reads the value array: no dictionary, or (where the format allows it) too many values
    const bool raw = !VS::on || (VS::raw_tiles && nd == 0);
    // The loads below run on indices clamped into the arrays (a tile has at least one row; the
    // column list and the dictionary carry one padding element) instead of under conditions: a
    // load under a condition compiles to its own branch that waits for every load in flight (per
    // lane or per wave alike), which turned the tile's loads into a chain of memory latencies.
    const int rl = r < r1 ? r : r0;
    int rb = ip[rl] - p0, re = ip[rl + 1] - p0;
    typename Op::Own own = op.own(rl);
    if (r >= r1) rb = re = 0;
    if (len > CAP)
    {
        // one row longer than the buffer: one lane, from global memory, in order
        typename Op::Part part{};
        if (!op.ready()) return part;
        if (threadIdx.x == 0)
        {
            T s = 0;
            for (int p = p0; p < p0 + len; ++p)
            {
                T a;
                if constexpr (VS::on) a = raw ? v[p] : vs.dval[d0 + vs.vidx[p]];
                else a = v[p];
                const T t = a * op.gather(ix[p]);
                s += t;
            }
            part = op.store(r0, s, own);
        }
        return part;
    }
    // 1. distinct columns (and the tile's value dictionary) and entry loads
    const int u0 = td.u0, nu = td.u1 - td.u0;
    // The first NUF slots of columns (and NGF groups of entries below) cover what a tile of finite-
    // element rows holds (P3: about 480 distinct columns per 2048 entries), so they load without a
    // condition; the rest only in the tile-wide branch of a tile that needs them.
    constexpr int NUF = (NU + 2) / 3;
    int uc[NU];
#pragma unroll
    for (int i = 0; i < NUF; ++i)
    {
        const int u = static_cast<int>(threadIdx.x) + i * THREADS;
        uc[i] = __builtin_nontemporal_load(ucol + u0 + (u < nu ? u : 0));
    }
    if (nu > NUF * THREADS)
    {
#pragma unroll
        for (int i = NUF; i < NU; ++i)
        {
            const int u = static_cast<int>(threadIdx.x) + i * THREADS;
            uc[i] = __builtin_nontemporal_load(ucol + u0 + (u < nu ? u : 0));
        }
    }
    T dvv = 0;
    if constexpr (VS::on) dvv = vs.dval[d0 + (static_cast<int>(threadIdx.x) < nd ? static_cast<int>(threadIdx.x) : 0)];
    const int a0 = p0 & ~3, e = p0 + len;
    Vals4<T> vv[NGRP];
    u32 di[NGRP];
    u16x4 cc[NGRP];
    // whole groups of four: every group up to the tile's end lies inside the matrix's last whole
    // group (groups past the end load the tile's first group again and are not used); only the tile
    // holding the matrix's final partial group, or one without entries, loads entry by entry
    const bool whole = len > 0 && int64_t(e) <= (nnz & ~int64_t(3));
    if (whole)
    {
        constexpr int NGF = CAP / (4 * THREADS) > 0 ? CAP / (4 * THREADS) : 1;
        auto group = [&](int g) {
            const int a = a0 + 4 * (static_cast<int>(threadIdx.x) + g * THREADS);
            const int ac = a < e ? a : a0;
            if (raw) vv[g].load(v, ac);
            else if constexpr (VS::on) di[g] = __builtin_nontemporal_load(reinterpret_cast<const u32 *>(vs.vidx + ac));
            cc[g] = __builtin_nontemporal_load(reinterpret_cast<const u16x4 *>(lidx + ac));
        };
#pragma unroll
        for (int g = 0; g < (NGF < NGRP ? NGF : NGRP); ++g) group(g);
        if (e > a0 + 4 * NGF * THREADS)
        {
#pragma unroll
            for (int g = NGF; g < NGRP; ++g) group(g);
        }
    }
    else
    {
#pragma unroll
        for (int g = 0; g < NGRP; ++g)
        {
            const int a = a0 + 4 * (static_cast<int>(threadIdx.x) + g * THREADS);
            if (a < e)
            {
                if (int64_t(a) + 4 <= nnz)
                {
                    if (raw) vv[g].load(v, a);
                    else if constexpr (VS::on) di[g] = __builtin_nontemporal_load(reinterpret_cast<const u32 *>(vs.vidx + a));
                    cc[g] = __builtin_nontemporal_load(reinterpret_cast<const u16x4 *>(lidx + a));
                }
                else
                {
                    if constexpr (VS::on) di[g] = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                    {
                        const bool in = int64_t(a) + j < nnz;
                        if (raw) vv[g].set(j, in ? v[a + j] : T(0));
                        else if constexpr (VS::on) di[g] |= (in ? u32(vs.vidx[a + j]) : 0u) << (8 * j);
                        cc[g][j] = in ? lidx[a + j] : static_cast<unsigned short>(0);
                    }
                }
            }
            else
            {
                cc[g] = u16x4{0, 0, 0, 0};
                if constexpr (VS::on) di[g] = 0;
            }
        }
    }
    // every load above is in flight; the op may still decide that the tile has nothing to do
    if (!op.ready()) return typename Op::Part{};
    T xg[NU];
#pragma unroll
    for (int i = 0; i < NUF; ++i) xg[i] = op.gather(uc[i]);     // past nu: a valid column, not stored
    if (nu > NUF * THREADS)
    {
#pragma unroll
        for (int i = NUF; i < NU; ++i) xg[i] = op.gather(uc[i]);
    }
#pragma unroll
    for (int i = 0; i < NU; ++i)
    {
        const int u = static_cast<int>(threadIdx.x) + i * THREADS;
        if (u < nu) xs[u] = xg[i];       // slots past NUF hold values only when nu reaches them
    }
    if constexpr (VS::on)
        if (static_cast<int>(threadIdx.x) < nd) vs.dv[threadIdx.x] = dvv;
    __syncthreads();
    // 2. products: every gather from LDS first (positions masked into the buffer: an entry outside
    //    the tile reads some slot and is not stored), then the stores of the tile's own entries
    static_assert((CAP & (CAP - 1)) == 0, "CAP is a power of two");
    T pv[NGRP][4];
#pragma unroll
    for (int g = 0; g < NGRP; ++g)
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            T val;
            if (raw) val = vv[g].get(j);
            else if constexpr (VS::on) val = vs.dv[(di[g] >> (8 * j)) & 0xffu];
            pv[g][j] = val * xs[cc[g][j] & (CAP - 1)];
        }
#pragma unroll
    for (int g = 0; g < NGRP; ++g)
    {
        const int a = a0 + 4 * (static_cast<int>(threadIdx.x) + g * THREADS);
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int k = a + j - p0;
            if (k >= 0 && k < len) prod[k] = pv[g][j];
        }
    }
    __syncthreads();
    // 3. row sums, left to right, four loads in flight at a time
    typename Op::Part part{};
    if (r < r1)
    {
        T s = 0;
        int k = rb;
        for (; k + 3 < re; k += 4)
        {
            const T t0 = prod[k], t1 = prod[k + 1], t2 = prod[k + 2], t3 = prod[k + 3];
            s += t0;
            s += t1;
            s += t2;
            s += t3;
        }
        for (; k < re; ++k) s += prod[k];
        part = op.store(r, s, own);
    }
    return part;
}


// One tile (workgroup blockIdx.x) of a plan's tile table: bounds from blk (row, entry) and uptr.
// `stop` (may be NULL) is read together with the tile's bounds: a set flag ends the workgroup
// before it loads the matrix, and its latency hides behind theirs.  With `mask` (may be NULL) the
// tile is skipped when mask[tile_sys[b]] is zero (tile b belongs to system tile_sys[b]).
// *stopped tells the caller.
template <typename T, int THREADS, int CAP, class Op, class VS = NoDict>
__device__ inline typename Op::Part stream_tile(const int2 *__restrict__ blk, const int *__restrict__ uptr, const int *__restrict__ ucol,
                                     const unsigned short *__restrict__ lidx, const int *__restrict__ ip,
                                     const int *__restrict__ ix, const T *__restrict__ v, int64_t nnz, const Op &op, T *xs,
                                     T *prod, const int *stop, bool *stopped, const int *tile_sys = nullptr,
                                     const int *mask = nullptr, const VS &vs = VS{})
{
    const int b = blockIdx.x;
    const int2 b0 = blk[b], b1 = blk[b + 1];
    const int halt = (stop && *stop) || (mask && !mask[tile_sys[b]]);
    *stopped = halt != 0;
    if (halt) return typename Op::Part{};
    const TileDesc td{b0.x, b0.y, uptr[b], b1.x, b1.y, uptr[b + 1]};
    return tile_work<T, THREADS, CAP>(td, b, ucol, lidx, ip, ix, v, nnz, op, xs, prod, vs);
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
    __device__ bool ready() const { return true; }
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
                                                       int64_t nnz, const int *stop, const int *tile_sys, const int *mask)
{
    __shared__ T xs[CAP];
    __shared__ T prod[CAP];
    bool stopped = false;
    (void)stream_tile<T, THREADS, CAP>(blk, uptr, ucol, lidx, ip, ix, v, nnz, PlainOp<T>{x, y}, xs, prod, stop, &stopped,
                                       tile_sys, mask);
}

// ------------------------------------------------------------------ product of the inner CG
// The CG product with the direction update folded in: p' = r + beta p is formed from the gathered
// (r, p) at every distinct column and for the workgroup's own rows, which it writes to the other
// direction buffer together with q = A p'; the workgroup's partial of p'.q is returned.
template <typename T> struct PCGOwn {
    T r = 0, p = 0;
};
template <typename T> struct PCGOp {
    using Own = PCGOwn<T>;
    using Part = double;
    const T *r, *pold;
    T *pnew, *q;
    T beta;
    int active;
    __device__ bool ready() const { return active != 0; }
    __device__ T gather(int j) const { return fma(beta, pold[j], r[j]); }
    __device__ Own own(int i) const
    {
        Own o;
        o.r = r[i];
        o.p = pold[i];
        return o;
    }
    __device__ Part store(int i, T s, Own o) const
    {
        const T pn = fma(beta, o.p, o.r);
        pnew[i] = pn;
        q[i] = s;
        return static_cast<double>(pn) * static_cast<double>(s);
    }
};

} // namespace
} // namespace femd_amd
