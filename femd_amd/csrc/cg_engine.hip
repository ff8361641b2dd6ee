// cg_engine.hip -- the inner CG engine (one matrix or a batch of systems, FP32 or FP64): scaled
// values, the four iteration kernels, the reliable-update maintenance and the host driver.
#include "cg_engine.hpp"
#include "kernels.hpp"
#include "tile.hpp"

#include <hip/hip_ext.h>

#include <cmath>
#include <cstring>
#include <mutex>

namespace femd_amd {
namespace {

constexpr int BT = 256;
constexpr int BCHUNK = 1024;
constexpr int RT = 256;            // threads of the per-system reductions
constexpr int GRAPH_ITERS = 16;
constexpr int64_t FUSE_ALPHA_MAX = int64_t(1) << 20;   // tiles x chunks of the largest system (8 MiB of extra reads)
constexpr int UPDATE_WGS = 512;                          // one large system: update workgroups reading the tile partials
                                                         // (C1: 408 ms per solve; 256: 425, 1024: 410)
constexpr int64_t FUSE_ALPHA_RUN_MAX = int64_t(1) << 23; // ... and their reads at most (64 MiB)

inline int grid1(int64_t n) { return static_cast<int>(std::max<int64_t>(1, (n + BT - 1) / BT)); }

// The W sums out[w] = sum of a[i * W + w] over i in [0, n), in a fixed order, returned to every
// thread of the workgroup: lane l adds its elements i = l + k THREADS in k order, loading K of
// them per round all at once (one memory latency per round, not per element), then the wave
// (shuffles) and the waves in order.  The order depends only on n and THREADS.  red holds
// W THREADS / 64 doubles.  COH: a is read with device-coherent loads (values other workgroups of
// the running kernel stored with device-scope stores; see arrive_last).
template <int THREADS, int W, int K = (W == 1 ? 16 : 8), bool COH = false>
__device__ inline void wg_fixed_sums(const double *a, int n, double *red, double (&out)[W])
{
    double s[W];
#pragma unroll
    for (int w = 0; w < W; ++w) s[w] = 0.0;
    for (int base = 0; base < n; base += K * THREADS)
    {
        double v[K][W];
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            const int i = base + k * THREADS + static_cast<int>(threadIdx.x);
#pragma unroll
            for (int w = 0; w < W; ++w)
            {
                const double *e = a + static_cast<int64_t>(i) * W + w;
                if constexpr (COH) v[k][w] = i < n ? __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
                else v[k][w] = i < n ? *e : 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int w = 0; w < W; ++w) s[w] += v[k][w];
    }
#pragma unroll
    for (int w = 0; w < W; ++w)
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s[w] += __shfl_xor(s[w], off, 64);
        if ((threadIdx.x & 63) == 0) red[W * (threadIdx.x >> 6) + w] = s[w];
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < W; ++w)
    {
        out[w] = 0.0;
        for (int q = 0; q < THREADS / 64; ++q) out[w] += red[W * q + w];
    }
    __syncthreads();
}

// Four consecutive values of a vector (16-byte aligned): one float4, or two double2.
template <typename T> struct Vec4;
template <> struct Vec4<float> {
    float v[4];
    __device__ void load(const float *p)
    {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
    __device__ void store(float *p) const { *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]); }
};
template <> struct Vec4<double> {
    double v[4];
    __device__ void load(const double *p)
    {
        const double2 a = *reinterpret_cast<const double2 *>(p), b = *reinterpret_cast<const double2 *>(p + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
    __device__ void store(double *p) const
    {
        *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
    }
};
static_assert(BCHUNK == 4 * BT, "a whole chunk is four rows per lane");

// ---- scaled values ---------------------------------------------------------------------------
// one thread per position k of the inner CG's order (row i = perm[k]; perm NULL: i = k):
// sc[i] = a_ii^(-1/2) in the original order, dg[k] = a_ii in the inner CG's
__global__ __launch_bounds__(BT) void k_scale_of_diag(int n, const int *__restrict__ perm, const int *diag, const double *v,
                                                      double *sc, float *dg, int *bad_row)
{
    const int k = blockIdx.x * BT + threadIdx.x;
    if (k >= n) return;
    const int i = perm ? perm[k] : k;
    const int d = diag[i];
    const double a = d >= 0 ? v[d] : 0.0;
    dg[k] = static_cast<float>(a);
    if (!(a > 0.0))
    {
        atomicMin(bad_row, i);
        sc[i] = 0.0;
    }
    else sc[i] = 1.0 / std::sqrt(a);
}

// S A S in the inner CG's order: entry p of the reordered pattern (ip, ix; row k) is entry emap[p]
// of A, in row perm[k] and column perm[ix[p]], so out[p] = v[emap[p]] s_i s_j -- the same product,
// in the same order, as for the entry in place (one lane per row: a set-up pass).  perm and emap
// NULL: the original order.
// Sixteen consecutive lanes per row, so that a row's entries are read and written together.
constexpr int SV_LANES = 16;
template <typename T>
__global__ __launch_bounds__(BT) void k_scaled_values(int n, const int *ip, const int *ix, const int *__restrict__ emap,
                                                      const int *__restrict__ perm, const double *v, const double *sc, T *out)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * BT + threadIdx.x;
    const int k = static_cast<int>(t / SV_LANES), j = static_cast<int>(t % SV_LANES);
    if (k >= n) return;
    const int p0 = ip[k], p1 = ip[k + 1];
    if (!perm)
    {
        const double si = sc[k];
        for (int p = p0 + j; p < p1; p += SV_LANES) out[p] = static_cast<T>(v[p] * si * sc[ix[p]]);
        return;
    }
    const double si = sc[perm[k]];
    for (int p = p0 + j; p < p1; p += SV_LANES) out[p] = static_cast<T>(v[emap[p]] * si * sc[perm[ix[p]]]);
}

// ---- iteration kernels -----------------------------------------------------------------------
// The product over the list of tiles of the running systems (k_cg_list_fill): entry g is two
// int4 -- the tile's first row, first entry, first distinct column and system, then the same
// bounds of the next tile and the tile's index -- so a tile starts from one load.  The direction
// buffers alternate with the iteration (odd: read P1, write P0), the same for every system, so
// the own-row and column loads do not wait for the system's state; a system that has paused since
// the list was built skips its tiles (PCGOp::ready) once those loads are in flight.
// DICT: 0 the value array, 1 per-tile dictionaries, 2 dictionaries where a tile has one (the
// others read the value array)
template <typename T, int THREADS, int CAP, int DICT>
__global__ __launch_bounds__(THREADS) void k_cg_product(const int4 *__restrict__ list, int count,
                                                        const int *__restrict__ ucol, const unsigned short *__restrict__ lidx,
                                                        const int *__restrict__ ip, const int *__restrict__ ix,
                                                        const T *__restrict__ v, const unsigned char *__restrict__ vidx,
                                                        const int *__restrict__ doff, const T *__restrict__ dval,
                                                        const T *__restrict__ r, T *P0, T *P1, T *__restrict__ q, int64_t nnz,
                                                        int odd, double *__restrict__ tparts, const CGSys *st)
{
    __shared__ T xs[CAP];
    __shared__ T prod[CAP];
    __shared__ T dv[DICT ? 256 : 1];
    __shared__ double red[THREADS / 64];
    const int g = blockIdx.x;
    if (g >= count) return;
    const int4 e0 = list[2 * g], e1 = list[2 * g + 1];
    const CGSys &S = st[e0.w];
    const PCGOp<T> op{r, odd ? P1 : P0, odd ? P0 : P1, q, static_cast<T>(S.beta), S.active};
    const TileDesc td{e0.x, e0.y, e0.z, e1.x, e1.y, e1.z};
    double part;
    if constexpr (DICT == 2)
        part = tile_work<T, THREADS, CAP>(td, e1.w, ucol, lidx, ip, ix, v, nnz, op, xs, prod,
                                          TileDict<T, true>{vidx, doff, dval, dv});
    else if constexpr (DICT == 1)
        part = tile_work<T, THREADS, CAP>(td, e1.w, ucol, lidx, ip, ix, v, nnz, op, xs, prod, TileDict<T>{vidx, doff, dval, dv});
    else
        part = tile_work<T, THREADS, CAP>(td, e1.w, ucol, lidx, ip, ix, v, nnz, op, xs, prod);
    if (!op.ready()) return;
    const double t = wg_sum<THREADS>(part, red);
    if (threadIdx.x == 0) tparts[e1.w] = t;
}

// one thread per tile: the list entry of every tile of a system with aoff[sys] >= 0, at
// aoff[sys] + its position among the system's tiles
__global__ __launch_bounds__(BT) void k_cg_list_fill(int ntiles, const int2 *__restrict__ blk, const int *__restrict__ uptr,
                                                     const int *__restrict__ tile_sys, const int *__restrict__ seg_tiles,
                                                     const int *__restrict__ aoff, int4 *list)
{
    const int b = blockIdx.x * BT + threadIdx.x;
    if (b >= ntiles) return;
    const int sys = tile_sys ? tile_sys[b] : 0;
    const int a = aoff[sys];
    if (a < 0) return;
    const int k = a + b - seg_tiles[sys];
    const int2 b0 = blk[b], b1 = blk[b + 1];
    list[2 * k] = make_int4(b0.x, b0.y, uptr[b], sys);
    list[2 * k + 1] = make_int4(b1.x, b1.y, uptr[b + 1], b);
}

// one workgroup per system: p'.q over its tiles, alpha.  The sum runs in the order of the update
// kernels' own (BT lanes), so alpha has the same bits whether it is formed here or there; 48 loads
// per lane per round keep a system of 12 k tiles to one round.
__global__ __launch_bounds__(BT) void k_cg_alpha(const int *__restrict__ seg_tiles, const double *__restrict__ tparts, CGSys *st)
{
    CGSys &S = st[blockIdx.x];
    if (!S.active) return;
    __shared__ double red[BT / 64];
    const int t0 = seg_tiles[blockIdx.x], t1 = seg_tiles[blockIdx.x + 1];
    double o[1];
    wg_fixed_sums<BT, 1, 48>(tparts + t0, t1 - t0, red, o);
    const double pq = o[0];
    if (threadIdx.x != 0) return;
    if (!(pq > 0.0))
    {
        S.active = 0;
        S.reason = 2;
        return;
    }
    S.alpha = S.rr / pq;
}

// The end of an iteration for one system (thread 0): the new r.r and r.D r, the pause tests, beta.
__device__ inline void beta_verdict(CGSys &S, double rr, double rt, int odd)
{
    S.it += 1;
    S.total += 1;
    S.dir = odd ^ 1;                        // the new direction's buffer
    int reason = 0;
    if (!(rr == rr) || !(rt == rt)) reason = 2;
    else if (rr <= S.thr || rt <= S.tgt) reason = 1;
    else if (S.it >= S.max_iter) reason = 3;
    if (reason)
    {
        S.active = 0;
        S.reason = reason;
        S.rr = rr;
        return;
    }
    S.beta = rr / S.rr;
    S.rr = rr;
}

// Arrival counters: the update workgroups of a system count themselves in, and the last one forms
// the system's beta, so that no kernel runs for it.  Same-address device-scope atomics serialise
// (about 10 ns each on MI355X: 1444 arrivals on one counter cost 15 us), so a system with more
// than 64 workgroups counts in two levels: workgroup w adds to first-level counter w % 64, the last
// arrival of each adds to the system's counter.  Counters lie 64 ints (256 bytes) apart.  Each
// chunk partial is written with a device-scope store (it writes through the XCD's L2) and the
// wave waits for those stores before its first atomic, so the last workgroup, reading with
// device-scope loads, sees every partial; no device-scope fence is needed (one writes back the
// whole L2 of the XCD, which costs more than the kernel it would save).  The last arrival of each
// counter clears it for the next iteration.  Valid in thread 0.
__device__ inline bool arrive_last(int *cnt, int nw, int w)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (nw <= 64)
    {
        if (__hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != nw - 1) return false;
        __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    const int g = w & 63, size = nw / 64 + (g < nw % 64 ? 1 : 0);
    int *c1 = cnt + 64 * (1 + g);
    if (__hip_atomic_fetch_add(c1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != size - 1) return false;
    __hip_atomic_store(c1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 63) return false;
    __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// y += alpha p', r -= alpha q over one chunk of one system, and the chunk's partials of r.r and
// r.D r (D = diag(A): rho^2 r.D r is the squared norm of the unscaled residual r stands for).
// With FUSE (small systems) every workgroup forms alpha = r.r / p'.q itself from the system's tile
// partials, in one fixed order (the same bits in every workgroup), and k_cg_alpha is not launched;
// the system's first chunk records alpha.  The system's last workgroup (arrive_last: counters at
// acnt + abase[sys], anw[sys] workgroups) forms beta from the chunk partials (BT lanes, the order
// of the other per-system sums), or pauses the system after a breakdown (p'.q <= 0, seen alike
// by every workgroup).  A workgroup holds the chunks of one system.
template <typename T, bool FUSE>
__global__ __launch_bounds__(BT) void k_cg_update(const int64_t *__restrict__ clo, const int *__restrict__ cpid, T *y, T *r,
                                                  const T *P0, const T *P1, const T *__restrict__ q,
                                                  const float *__restrict__ dg, int odd, double *cparts,
                                                  const int *__restrict__ seg_tiles, const int *__restrict__ cptr,
                                                  const double *tparts, CGSys *st, int nch, int per, int *acnt,
                                                  const int *__restrict__ abase, const int *__restrict__ anw)
{
    __shared__ double red[2 * (BT / 64)];
    __shared__ int last;
    const T *pnew = odd ? P0 : P1;
    // chunks [c0, c1) of this workgroup (one, or with FUSE for one large system a short run, so that
    // fewer workgroups read the tile partials)
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, nch);
    int alpha_sys = -1, run_sys = -1;
    bool broken = false;
    double alpha = 0.0;
    for (int c = c0; c < c1; ++c)
    {
        const int sys = cpid[c];
        CGSys &S = st[sys];
        if (!S.active) continue;
        run_sys = sys;
        if constexpr (FUSE)
        {
            if (sys != alpha_sys)
            {
                double o[1];
                const int t0 = seg_tiles[sys];
                wg_fixed_sums<BT, 1>(tparts + t0, seg_tiles[sys + 1] - t0, red, o);
                alpha_sys = sys;
                if (!(o[0] > 0.0))
                {
                    broken = true;
                    break;
                }
                alpha = S.rr / o[0];
            }
            if (c == cptr[sys] && threadIdx.x == 0) S.alpha = alpha;
        }
        else
            alpha = S.alpha;
        const T a = static_cast<T>(alpha);
        double s = 0.0, t = 0.0;
        const int64_t lo = clo[c], hi = clo[c + 1];
        if ((lo & 3) == 0 && hi - lo == BCHUNK)
        {
            // a whole aligned chunk: four consecutive rows per lane, vector loads and stores
            const int64_t i = lo + 4 * static_cast<int64_t>(threadIdx.x);
            Vec4<T> yv, pv, rv, qv;
            yv.load(y + i); pv.load(pnew + i); rv.load(r + i); qv.load(q + i);
            const float4 dv = *reinterpret_cast<const float4 *>(dg + i);
            const float dd[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                yv.v[j] = fma(a, pv.v[j], yv.v[j]);
                const T ri = fma(-a, qv.v[j], rv.v[j]);
                rv.v[j] = ri;
                const double e = static_cast<double>(ri) * static_cast<double>(ri);
                s += e;
                t += e * static_cast<double>(dd[j]);
            }
            yv.store(y + i);
            rv.store(r + i);
        }
        else
            for (int64_t i = lo + threadIdx.x; i < hi; i += BT)
            {
                y[i] = fma(a, pnew[i], y[i]);
                const T ri = fma(-a, q[i], r[i]);
                r[i] = ri;
                const double e = static_cast<double>(ri) * static_cast<double>(ri);
                s += e;
                t += e * static_cast<double>(dg[i]);
            }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
        {
            s += __shfl_xor(s, off, 64);
            t += __shfl_xor(t, off, 64);
        }
        if ((threadIdx.x & 63) == 0)
        {
            red[2 * (threadIdx.x >> 6)] = s;
            red[2 * (threadIdx.x >> 6) + 1] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double a0 = 0.0, a1 = 0.0;
            for (int w = 0; w < BT / 64; ++w)
            {
                a0 += red[2 * w];
                a1 += red[2 * w + 1];
            }
            __hip_atomic_store(cparts + 2 * c, a0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(cparts + 2 * c + 1, a1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();                      // red is reused by the next chunk
    }
    // the system's last workgroup: beta, or the pause after a breakdown
    if (run_sys < 0) return;                  // the system is not running: nothing to count
    const int sys = run_sys;
    if (threadIdx.x == 0)
        last = arrive_last(acnt + abase[sys], anw[sys], per > 1 ? static_cast<int>(blockIdx.x) : c0 - cptr[sys]);
    __syncthreads();
    if (!last) return;
    CGSys &S = st[sys];
    if (broken)
    {
        if (threadIdx.x == 0)
        {
            S.active = 0;
            S.reason = 2;
        }
        return;
    }
    const int k0 = cptr[sys], k1 = cptr[sys + 1];
    double o[2];
    wg_fixed_sums<BT, 2, 8, true>(cparts + 2 * k0, k1 - k0, red, o);
    if (threadIdx.x == 0) beta_verdict(S, o[0], o[1], odd);
}

// ---- maintenance kernels (on the systems with sel[sys] != 0) ---------------------------------
// x (original order) += rho S y (y in the inner CG's order: position i is unknown perm[i], or i
// when perm is NULL)
template <typename T>
__global__ __launch_bounds__(BT) void k_cg_finish(const int64_t *__restrict__ clo, const int *__restrict__ cpid,
                                                  const int *__restrict__ sel, const int *__restrict__ perm,
                                                  const double *__restrict__ sc, const T *__restrict__ y, const CGSys *st,
                                                  double *x)
{
    const int c = blockIdx.x, sys = cpid[c];
    if (!sel[sys]) return;
    const double rho = st[sys].rho;
    if (!(rho > 0.0)) return;
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT)
    {
        const int64_t j = perm ? perm[i] : i;
        x[j] += rho * (sc[j] * static_cast<double>(y[i]));
    }
}

// r64 = b - t64 and the chunk's partial of r64.r64 (of b.b when t64 is NULL)
__global__ __launch_bounds__(BT) void k_cg_residual(const int64_t *__restrict__ clo, const int *__restrict__ cpid,
                                                    const int *__restrict__ sel, const double *__restrict__ b,
                                                    const double *__restrict__ t64, double *r64, double *cparts)
{
    __shared__ double red[BT / 64];
    const int c = blockIdx.x, sys = cpid[c];
    if (!sel[sys]) return;
    double s = 0.0;
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT)
    {
        const double ri = t64 ? b[i] - t64[i] : b[i];
        if (t64) r64[i] = ri;
        s += ri * ri;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0.0;
        for (int w = 0; w < BT / 64; ++w) t += red[w];
        cparts[c] = t;
    }
}

// one workgroup per system: the sum of its chunks' partials
__global__ __launch_bounds__(RT) void k_cg_sys_sum(const int *__restrict__ cptr, const int *__restrict__ sel,
                                                   const double *__restrict__ cparts, double *out)
{
    if (!sel[blockIdx.x]) return;
    __shared__ double red[RT / 64];
    const int c0 = cptr[blockIdx.x], c1 = cptr[blockIdx.x + 1];
    double o[1];
    wg_fixed_sums<RT, 1>(cparts + c0, c1 - c0, red, o);
    const double t = o[0];
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Resume (sel 1) or restart with no direction (sel 2): r = S r64 / rho_new (r64 in the original
// order, r in the inner CG's), y = 0, and the current direction p and q = A p scaled by
// rho / rho_new (or zeroed).
template <typename T>
__global__ __launch_bounds__(BT) void k_cg_rescale(const int64_t *__restrict__ clo, const int *__restrict__ cpid,
                                                   const int *__restrict__ sel, const double *__restrict__ rho_new,
                                                   const int *__restrict__ perm, const double *__restrict__ sc,
                                                   const double *__restrict__ r64, T *r, T *y, T *P0, T *P1, T *q, int next,
                                                   const CGSys *st)
{
    const int c = blockIdx.x, sys = cpid[c];
    const int mode = sel[sys];
    if (!mode) return;
    const double rn = rho_new[sys];
    const double scale = rn > 0.0 ? 1.0 / rn : 0.0;
    const T f = static_cast<T>(mode == 2 || !(rn > 0.0) ? 0.0 : st[sys].rho / rn);
    const T *pc = st[sys].dir ? P1 : P0;    // the current direction ...
    T *pn = next ? P1 : P0;                  // ... moves to where the next product reads it
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT)
    {
        const int64_t j = perm ? perm[i] : i;
        r[i] = static_cast<T>(sc[j] * r64[j] * scale);
        y[i] = T(0);
        // a restart stores zeros (the buffers may hold anything, NaN included, before the first)
        if (mode == 2)
        {
            P0[i] = T(0);
            P1[i] = T(0);
            q[i] = T(0);
        }
        else
        {
            pn[i] = pc[i] * f;
            q[i] *= f;
        }
    }
}

// per chunk: the partials of r.q, p.q and r.r (p the current direction)
template <typename T>
__global__ __launch_bounds__(BT) void k_cg_fresh_dots(const int64_t *__restrict__ clo, const int *__restrict__ cpid,
                                                      const int *__restrict__ sel, const T *__restrict__ r, const T *P0,
                                                      const T *P1, const T *__restrict__ q, int next, double *c3)
{
    __shared__ double red[3 * (BT / 64)];
    const int c = blockIdx.x, sys = cpid[c];
    if (!sel[sys]) return;
    const T *pc = next ? P1 : P0;
    double a = 0.0, b = 0.0, d = 0.0;
    for (int64_t i = clo[c] + threadIdx.x; i < clo[c + 1]; i += BT)
    {
        const double ri = r[i], qi = q[i];
        a += ri * qi;
        b += static_cast<double>(pc[i]) * qi;
        d += ri * ri;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        d += __shfl_xor(d, off, 64);
    }
    if ((threadIdx.x & 63) == 0)
    {
        red[3 * (threadIdx.x >> 6)] = a;
        red[3 * (threadIdx.x >> 6) + 1] = b;
        red[3 * (threadIdx.x >> 6) + 2] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t[3] = {0.0, 0.0, 0.0};
        for (int w = 0; w < BT / 64; ++w)
            for (int j = 0; j < 3; ++j) t[j] += red[3 * w + j];
        for (int j = 0; j < 3; ++j) c3[3 * c + j] = t[j];
    }
}

// one workgroup per system: r.q, p.q, r.r over its chunks; beta that makes the next direction
// A-conjugate to the current one (0 on a restart), the segment's threshold, running again
__global__ __launch_bounds__(RT) void k_cg_fresh_state(const int *__restrict__ cptr, const int *__restrict__ sel,
                                                       const double *__restrict__ c3, const double *__restrict__ rho_new,
                                                       const double *__restrict__ tgt, double tol2, int max_iter, int next,
                                                       CGSys *st)
{
    const int sys = blockIdx.x, mode = sel[sys];
    if (!mode) return;
    __shared__ double red[3 * (RT / 64)];
    const int c0 = cptr[sys], c1 = cptr[sys + 1];
    double o[3];
    wg_fixed_sums<RT, 3>(c3 + 3 * c0, c1 - c0, red, o);
    const double rq = o[0], pq = o[1], rr = o[2];
    if (threadIdx.x != 0) return;
    CGSys &S = st[sys];
    const double rn = rho_new[sys];
    if (mode == 2)
    {
        S.it = 0;
        S.beta = 0.0;
    }
    else S.beta = pq > 0.0 ? -rq / pq : 0.0;
    S.max_iter = S.it + max_iter;           // a cap per segment
    S.dir = next;
    S.tol2 = tol2;
    S.tgt = tgt ? tgt[sys] : 0.0;
    S.rr = rr;
    S.thr = tol2 * rr;
    S.rho = rn;
    S.reason = 0;
    S.active = (rn > 0.0 && rr > 0.0) ? 1 : 0;
}

__global__ void k_cg_count(int np, const CGSys *st, const int *__restrict__ retired, int *out)
{
    __shared__ int cnt[2];
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    int run = 0, paused = 0;
    for (int s = threadIdx.x; s < np; s += blockDim.x)
    {
        if (st[s].active) ++run;
        else if (!retired[s] && st[s].reason) ++paused;
    }
    atomicAdd(&cnt[0], run);
    atomicAdd(&cnt[1], paused);
    __syncthreads();
    if (threadIdx.x == 0) { out[0] = cnt[0]; out[1] = cnt[1]; }
}

// ---- sampled timing of the product kernel ------------------------------------------------------
struct Sample {
    hipEvent_t e0, e1;
};
struct InnerTimer {
    std::mutex mu;
    bool on = false;
    int64_t launches = 0;
    std::vector<Sample> samples;
    size_t used = 0;
    std::string shape;             // launch shape of the last timed launch
};
InnerTimer g_timer;

// Events for this product launch, or false (timer off, every 8th launch only, never in capture).
bool timer_sample(hipStream_t s, hipEvent_t *e0, hipEvent_t *e1, const char *shape)
{
    std::lock_guard<std::mutex> lk(g_timer.mu);
    if (!g_timer.on) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    if ((g_timer.launches++ & 7) != 0) return false;
    if (g_timer.used == g_timer.samples.size())
    {
        Sample x{nullptr, nullptr};
        HIPC(hipEventCreate(&x.e0));
        HIPC(hipEventCreate(&x.e1));
        g_timer.samples.push_back(x);
    }
    const Sample &x = g_timer.samples[g_timer.used++];
    g_timer.shape = shape;
    *e0 = x.e0;
    *e1 = x.e1;
    return true;
}

bool timer_on()
{
    std::lock_guard<std::mutex> lk(g_timer.mu);
    return g_timer.on;
}

} // namespace

// ---- the inner CG's order (cached on the pattern) --------------------------------------------
const LocalOrder &local_order(DevPattern &P, hipStream_t s)
{
    if (P.local) return *P.local;
    if (P.n != P.m) fail(FEMD_EINVAL, "inner CG: the matrix must be square");
    const int n = P.n;
    if (device_setup_enabled())
    {
        auto L = std::make_unique<LocalOrder>();
        if (!device_first_met_order(n, P.ip.as<int>(), P.ix.as<int>(), P.nnz, P.segments, *L, s))
            fail(FEMD_EINVAL, "inner CG: a system's rows have columns outside it");
        P.local = std::move(L);
        return *P.local;
    }
    std::vector<int> perm, ip2, ix2, emap;
    if (!first_met_order(n, P.h_ip.data(), P.h_ix.data(), P.segments, perm, ip2, ix2, emap))
        fail(FEMD_EINVAL, "inner CG: a system's rows have columns outside it");
    perm.resize(size_t(std::max(n, 1)));
    emap.resize(size_t(std::max<int64_t>(P.nnz, 1)));
    auto L = std::make_unique<LocalOrder>();
    L->perm = upload(perm.data(), size_t(n), s);
    L->emap = upload(emap.data(), size_t(P.nnz), s);
    L->pat = make_pattern(n, n, std::move(ip2), std::move(ix2), s);
    L->pat->segments = P.segments;
    HIPC(hipStreamSynchronize(s));            // perm, emap go out of scope
    P.local = std::move(L);
    return *P.local;
}

// The LocalOrder the inner CG uses for P, or NULL (femd_inner_order off: the original order).
const LocalOrder *inner_order(DevPattern &P, hipStream_t s)
{
    return inner_order_enabled() ? &local_order(P, s) : nullptr;
}

// The launch plan of the inner CG's pattern Q (P itself, or its LocalOrder's pattern): the shape
// femd_spmv_select set on P, or Q's own tuned one.
const SpmvPlan &local_plan(DevPattern &P, DevPattern &Q, int prec, const void *values, hipStream_t s)
{
    if (&Q == &P) return spmv_plan(P, prec, values, s);
    const int f = P.forced[prec == FEMD_FP32 ? 1 : 0];
    std::unique_ptr<SpmvPlan> &slot = prec == FEMD_FP32 ? Q.plan32 : Q.plan64;
    if (f >= 0 && (!slot || slot->variant != f)) slot = std::make_unique<SpmvPlan>(shape_plan(Q, f, prec, s));
    return spmv_plan(Q, prec, values, s);
}

// ---- scaled values (cached on the matrix) --------------------------------------------------
void scaled_values(femd_csr_s *A, hipStream_t s, bool reorder)
{
    DevPattern &P = *A->pat;
    const LocalOrder *L = reorder ? inner_order(P, s) : nullptr;
    if (A->v32s_valid && A->scaled_reordered == (L != nullptr)) return;
    A->v64s_valid = false;
    if (P.n != P.m) fail(FEMD_EINVAL, "inner CG: the matrix must be square");
    const int n = P.n;
    const int *diag = diag_of(P, s);
    const DevPattern &Q = L ? *L->pat : P;
    const int *perm = L ? L->perm.as<int>() : nullptr, *emap = L ? L->emap.as<int>() : nullptr;
    A->scale.ensure(sizeof(double) * std::max(n, 1));
    A->diagf.ensure(sizeof(float) * std::max(n, 1));
    A->v32s.ensure(sizeof(float) * std::max<int64_t>(P.nnz, 1));
    DevBuf bad(sizeof(int));
    const int big = 0x7fffffff;
    HIPC(hipMemcpyAsync(bad.p, &big, sizeof(int), hipMemcpyHostToDevice, s));
    if (n > 0)
    {
        hipLaunchKernelGGL(k_scale_of_diag, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, perm, diag, A->v.as<double>(),
                           A->scale.as<double>(), A->diagf.as<float>(), bad.as<int>());
        hipLaunchKernelGGL(k_scaled_values<float>, dim3((int64_t(n) * SV_LANES + BT - 1) / BT), dim3(BT), 0, s, n, Q.ip.as<int>(), Q.ix.as<int>(),
                           emap, perm, A->v.as<double>(), A->scale.as<double>(), A->v32s.as<float>());
    }
    int badrow = big;
    HIPC(hipMemcpyAsync(&badrow, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (badrow != big)
        fail(FEMD_ENOTSPD, "inner CG: diagonal entry of row " + std::to_string(badrow) + " is not positive");
    A->v32s_valid = true;
    A->scaled_reordered = L != nullptr;
    A->d32s.valid = false;
}

void scaled_values64(femd_csr_s *A, hipStream_t s, bool reorder)
{
    scaled_values(A, s, reorder);
    if (A->v64s_valid) return;
    const LocalOrder *L = A->scaled_reordered ? &local_order(*A->pat, s) : nullptr;
    const DevPattern &Q = L ? *L->pat : *A->pat;
    A->v64s.ensure(sizeof(double) * std::max<int64_t>(Q.nnz, 1));
    if (Q.n > 0)
        hipLaunchKernelGGL(k_scaled_values<double>, dim3((int64_t(Q.n) * SV_LANES + BT - 1) / BT), dim3(BT), 0, s, Q.n, Q.ip.as<int>(),
                           Q.ix.as<int>(), L ? L->emap.as<int>() : nullptr, L ? L->perm.as<int>() : nullptr,
                           A->v.as<double>(), A->scale.as<double>(), A->v64s.as<double>());
    A->v64s_valid = true;
    A->d64s.valid = false;
}

// ---- the engine ------------------------------------------------------------------------------
CGEngine::~CGEngine() { drop_graph(); }

void CGEngine::drop_graph()
{
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    graph_stream = nullptr;
}

void CGEngine::setup(femd_csr_s *M, int prec, hipStream_t s, bool reorder)
{
    if (prec != FEMD_FP32 && prec != FEMD_FP64) fail(FEMD_EINVAL, "inner CG: precision is FEMD_FP32 or FEMD_FP64");
    if (M->pat->n != M->pat->m) fail(FEMD_EINVAL, "inner CG: the matrix must be square");
    if (prec == FEMD_FP64) scaled_values64(M, s, reorder);
    else scaled_values(M, s, reorder);
    // the plan of the pattern in the inner CG's order (femd_spmv_select may have replaced it, or
    // femd_inner_order switched the order, since the last bind)
    const LocalOrder *L = M->scaled_reordered ? &local_order(*M->pat, s) : nullptr;
    const SpmvPlan &want = local_plan(*M->pat, L ? *L->pat : *M->pat, prec, prec == FEMD_FP64 ? M->v64s.p : M->v32s.p, s);
    if (!(A == M && precision == prec && plan && plan_id == want.id))
    {
        drop_graph();
        A = M;
        precision = prec;
        bind(M, prec, s);
    }
    // the value dictionary of the scaled values under this plan, when the values allow one
    ValueDict &d = prec == FEMD_FP64 ? M->d64s : M->d32s;
    const bool want_dict = dict_enabled();
    if (want_dict && (!d.valid || d.plan_id != plan_id))
        build_dict(*plan, prec == FEMD_FP64 ? M->v64s.p : M->v32s.p, prec, M->pat->nnz, d, s);
    const bool use = want_dict && d.valid && d.usable && d.plan_id == plan_id;
    // a captured graph holds the format and the dictionary's buffers
    if (use != use_dict || (use && d.gen != dict_gen)) drop_graph();
    use_dict = use;
    dict_raw = use && d.raw_tiles > 0;
    dict_gen = use ? d.gen : 0;
}

void CGEngine::bind(femd_csr_s *M, int prec, hipStream_t s)
{
    const LocalOrder *L = M->scaled_reordered ? &local_order(*M->pat, s) : nullptr;
    DevPattern &P = L ? *L->pat : *M->pat;
    pat = &P;
    perm = L ? L->perm.as<int>() : nullptr;
    n = P.n;
    roff = P.segments.empty() ? std::vector<int64_t>{0, int64_t(P.n)} : P.segments;
    np = static_cast<int>(roff.size()) - 1;
    std::vector<int64_t> lo{0};
    std::vector<int> cp, cptr_h{0};
    max_rows = 0;
    for (int p = 0; p < np; ++p)
    {
        const int64_t r0 = roff[p], nr = roff[p + 1] - roff[p];
        max_rows = std::max(max_rows, nr);
        for (int64_t k = 0; k < nr; k += BCHUNK)
        {
            lo.push_back(r0 + std::min<int64_t>(k + BCHUNK, nr));
            cp.push_back(p);
        }
        cptr_h.push_back(static_cast<int>(cp.size()));
    }
    nch = static_cast<int>(cp.size());
    h_cptr = cptr_h;
    clo = upload(lo, s);
    cpid = upload(cp.empty() ? std::vector<int>{0} : cp, s);
    cptr = upload(cptr_h, s);
    plan = &local_plan(*M->pat, P, prec, prec == FEMD_FP64 ? M->v64s.p : M->v32s.p, s);
    plan_id = plan->id;
    seg_tiles = upload(plan->seg_tiles, s);
    // alpha folded into the update when every workgroup's extra reads stay small: each of a
    // system's chunks reads all of its tile partials
    int64_t most = 0;
    for (int p = 0; p < np; ++p)
        most = std::max(most, int64_t(plan->seg_tiles[size_t(p) + 1] - plan->seg_tiles[size_t(p)]) *
                                  int64_t(cptr_h[size_t(p) + 1] - cptr_h[size_t(p)]));
    fuse_alpha = most <= FUSE_ALPHA_MAX;
    update_per = 1;
    if (!fuse_alpha && np == 1)
    {
        // one large system: runs of chunks per workgroup, so that at most UPDATE_WGS workgroups
        // read the tile partials
        const int tiles = plan->seg_tiles[1] - plan->seg_tiles[0];
        const int per = (nch + UPDATE_WGS - 1) / UPDATE_WGS;
        if (int64_t(tiles) * ((nch + per - 1) / per) <= FUSE_ALPHA_RUN_MAX)
        {
            fuse_alpha = true;
            update_per = per;
        }
    }
    // arrival counters of the update workgroups (k_cg_update, arrive_last): per system one counter,
    // or one and 64 first-level ones when more than 64 workgroups hold its chunks; 64 ints apart
    {
        std::vector<int> base(size_t(std::max(np, 1)), 0), nwg(size_t(std::max(np, 1)), 0);
        int64_t at = 0;
        for (int p = 0; p < np; ++p)
        {
            const int nw = update_per > 1 ? (nch + update_per - 1) / update_per : cptr_h[size_t(p) + 1] - cptr_h[size_t(p)];
            nwg[size_t(p)] = nw;
            base[size_t(p)] = static_cast<int>(at);
            at += 64 * (nw > 64 ? 65 : 1);
        }
        acnt = DevBuf(4 * size_t(std::max<int64_t>(at, 64)));
        HIPC(hipMemsetAsync(acnt.p, 0, 4 * size_t(std::max<int64_t>(at, 64)), s));
        abase = upload(base, s);
        anw = upload(nwg, s);
    }
    const size_t N = size_t(std::max<int64_t>(n, 1)), w = prec == FEMD_FP64 ? 8 : 4;
    st = DevBuf(sizeof(CGSys) * size_t(std::max(np, 1)));
    tparts = DevBuf(8 * size_t(std::max(plan->ntiles, 1)));
    cparts = DevBuf(24 * size_t(std::max(nch, 1)));
    r = DevBuf(w * N); q = DevBuf(w * N); P0 = DevBuf(w * N); P1 = DevBuf(w * N); y = DevBuf(w * N);
    r64 = DevBuf(8 * N); t64 = DevBuf(8 * N);
    sel = DevBuf(4 * size_t(std::max(np, 1))); retired = DevBuf(4 * size_t(std::max(np, 1)));
    rho_new = DevBuf(8 * size_t(std::max(np, 1)));
    tgt = DevBuf(8 * size_t(std::max(np, 1)));
    cnt = DevBuf(2 * sizeof(int));
    list = DevBuf(32 * size_t(std::max(plan->ntiles, 1)));
    aoff = DevBuf(4 * size_t(std::max(np, 1)));
    parity = 0;
    list_count = -1;
    build_list(nullptr, s);
    std::vector<CGSys> z(size_t(std::max(np, 1)));
    std::memset(z.data(), 0, sizeof(CGSys) * z.size());
    HIPC(hipMemcpyAsync(st.p, z.data(), sizeof(CGSys) * z.size(), hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
}

// The product's tile list: the tiles of every system not marked in `leave_out` (NULL: all).
void CGEngine::build_list(const std::vector<int> *leave_out, hipStream_t s)
{
    const std::vector<int> &stl = plan->seg_tiles;
    std::vector<int> off(size_t(std::max(np, 1)), -1);
    int run = 0;
    for (int p = 0; p < np; ++p)
    {
        if (leave_out && (*leave_out)[size_t(p)]) continue;
        off[size_t(p)] = run;
        run += stl[size_t(p) + 1] - stl[size_t(p)];
    }
    HIPC(hipMemcpyAsync(aoff.p, off.data(), 4 * off.size(), hipMemcpyHostToDevice, s));
    if (plan->ntiles > 0)
        hipLaunchKernelGGL(k_cg_list_fill, dim3((plan->ntiles + BT - 1) / BT), dim3(BT), 0, s, plan->ntiles,
                           plan->blocks.as<int2>(), plan->uptr.as<int>(), plan->tile_sys.p ? plan->tile_sys.as<int>() : nullptr,
                           seg_tiles.as<int>(), aoff.as<int>(), list.as<int4>());
    HIPC(hipStreamSynchronize(s));            // off goes out of scope
    if (run != list_count) drop_graph();      // a captured graph holds the launch's grid
    list_count = run;
}

namespace {

template <typename T, int DICT>
void launch_product(CGEngine &E, int odd, hipStream_t s)
{
    const DevPattern &P = *E.pat;
    const SpmvPlan &pl = *E.plan;
    int th = 0, cap = 0;
    plan_shape(pl, &th, &cap);
    const int *ucol = pl.ucol.as<int>();
    const unsigned short *lidx = pl.lidx.as<unsigned short>();
    const T *v = sizeof(T) == 8 ? reinterpret_cast<const T *>(E.A->v64s.p) : reinterpret_cast<const T *>(E.A->v32s.p);
    const ValueDict &d = sizeof(T) == 8 ? E.A->d64s : E.A->d32s;
    const unsigned char *vidx = DICT ? d.vidx.as<unsigned char>() : nullptr;
    const int *doff = DICT ? d.doff.as<int>() : nullptr;
    const T *dval = DICT ? d.dval.as<T>() : nullptr;
    const CGSys *st = E.st.as<CGSys>();
    const int4 *list = E.list.as<int4>();
    const int count = E.list_count, grid = std::max(count, 1);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!timer_sample(s, &e0, &e1, spmv_variant_name(pl))) e0 = e1 = nullptr;
    bool launched = false;
#define FEMD_CASE(TH, CP)                                                                                             \
    if constexpr (2 * CP * sizeof(T) + (DICT ? 256 * sizeof(T) : 0) <= 65536)                                        \
        if (th == TH && cap == CP)                                                                                    \
        {                                                                                                             \
            launched = true;                                                                                          \
            if (e0)                                                                                                   \
                hipExtLaunchKernelGGL((k_cg_product<T, TH, CP, DICT>), dim3(grid), dim3(TH), 0, s, e0, e1, 0, list, count, \
                                      ucol, lidx, P.ip.as<int>(), P.ix.as<int>(), v, vidx, doff, dval, E.r.as<T>(),   \
                                      E.P0.as<T>(), E.P1.as<T>(), E.q.as<T>(), P.nnz, odd, E.tparts.as<double>(), st); \
            else                                                                                                      \
                hipLaunchKernelGGL((k_cg_product<T, TH, CP, DICT>), dim3(grid), dim3(TH), 0, s, list, count, ucol, lidx,  \
                                   P.ip.as<int>(), P.ix.as<int>(), v, vidx, doff, dval, E.r.as<T>(), E.P0.as<T>(),     \
                                   E.P1.as<T>(), E.q.as<T>(), P.nnz, odd, E.tparts.as<double>(), st);                 \
        }
    FEMD_SHAPES(FEMD_CASE)
#undef FEMD_CASE
    if (!launched) fail(FEMD_EINVAL, "inner CG: no product kernel for the plan's launch shape");
}

template <typename T>
void iterate_t(CGEngine &E, int count, hipStream_t s)
{
    CGSys *st = E.st.as<CGSys>();
    for (int k = 0; k < count; ++k)
    {
        const int odd = E.parity;
        if (E.use_dict && E.dict_raw) launch_product<T, 2>(E, odd, s);
        else if (E.use_dict) launch_product<T, 1>(E, odd, s);
        else launch_product<T, 0>(E, odd, s);
        int *ac = E.acnt.as<int>();
        const int *ab = E.abase.as<int>(), *an = E.anw.as<int>();
        if (E.fuse_alpha)
        {
            const int per = E.update_per;
            hipLaunchKernelGGL((k_cg_update<T, true>), dim3((E.nch + per - 1) / per), dim3(BT), 0, s, E.clo.as<int64_t>(),
                               E.cpid.as<int>(), E.y.as<T>(), E.r.as<T>(), E.P0.as<T>(), E.P1.as<T>(), E.q.as<T>(),
                               E.A->diagf.as<float>(), odd, E.cparts.as<double>(), E.seg_tiles.as<int>(), E.cptr.as<int>(),
                               E.tparts.as<double>(), st, E.nch, per, ac, ab, an);
        }
        else
        {
            hipLaunchKernelGGL(k_cg_alpha, dim3(E.np), dim3(BT), 0, s, E.seg_tiles.as<int>(), E.tparts.as<double>(), st);
            hipLaunchKernelGGL((k_cg_update<T, false>), dim3(E.nch), dim3(BT), 0, s, E.clo.as<int64_t>(), E.cpid.as<int>(),
                               E.y.as<T>(), E.r.as<T>(), E.P0.as<T>(), E.P1.as<T>(), E.q.as<T>(), E.A->diagf.as<float>(), odd,
                               E.cparts.as<double>(), E.seg_tiles.as<int>(), E.cptr.as<int>(), E.tparts.as<double>(), st,
                               E.nch, 1, ac, ab, an);
        }
        E.parity ^= 1;
    }
}

template <typename T>
void resume_t(CGEngine &E, const double *r64_dev, const double *rho_dev, const double *tgt_dev, const int *sel_dev,
              double tol2, int max_iter, hipStream_t s)
{
    const int next = E.parity;
    hipLaunchKernelGGL(k_cg_rescale<T>, dim3(E.nch), dim3(BT), 0, s, E.clo.as<int64_t>(), E.cpid.as<int>(), sel_dev, rho_dev,
                       E.perm, E.A->scale.as<double>(), r64_dev, E.r.as<T>(), E.y.as<T>(), E.P0.as<T>(), E.P1.as<T>(), E.q.as<T>(),
                       next, static_cast<const CGSys *>(E.st.as<CGSys>()));
    hipLaunchKernelGGL(k_cg_fresh_dots<T>, dim3(E.nch), dim3(BT), 0, s, E.clo.as<int64_t>(), E.cpid.as<int>(), sel_dev,
                       E.r.as<T>(), E.P0.as<T>(), E.P1.as<T>(), E.q.as<T>(), next, E.cparts.as<double>());
    hipLaunchKernelGGL(k_cg_fresh_state, dim3(E.np), dim3(RT), 0, s, E.cptr.as<int>(), sel_dev, E.cparts.as<double>(), rho_dev,
                       tgt_dev, tol2, max_iter, next, E.st.as<CGSys>());
}

template <typename T>
void finish_t(CGEngine &E, const int *sel_dev, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_finish<T>, dim3(E.nch), dim3(BT), 0, s, E.clo.as<int64_t>(), E.cpid.as<int>(), sel_dev,
                       E.perm, E.A->scale.as<double>(), E.y.as<T>(), static_cast<const CGSys *>(E.st.as<CGSys>()), x);
}

} // namespace

void CGEngine::iterate(int count, hipStream_t s)
{
    if (precision == FEMD_FP64) iterate_t<double>(*this, count, s);
    else iterate_t<float>(*this, count, s);
}

void CGEngine::replay(bool use_graph, hipStream_t s)
{
    if (!use_graph)
    {
        iterate(GRAPH_ITERS, s);
        return;
    }
    if (!exec || graph_stream != s)
    {
        drop_graph();
        HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        iterate(GRAPH_ITERS, s);
        HIPC(hipStreamEndCapture(s, &graph));
        HIPC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        graph_stream = s;
    }
    HIPC(hipGraphLaunch(exec, s));
}

void CGEngine::resume(const double *r64_dev, const double *rho_dev, const double *tgt_dev, const int *sel_dev, double tol2,
                      int max_iter, hipStream_t s)
{
    if (precision == FEMD_FP64) resume_t<double>(*this, r64_dev, rho_dev, tgt_dev, sel_dev, tol2, max_iter, s);
    else resume_t<float>(*this, r64_dev, rho_dev, tgt_dev, sel_dev, tol2, max_iter, s);
}

void CGEngine::start_device(const double *r64_dev, const double *rho_dev, const double *tgt_dev, const int *sel_dev,
                            double inner_tol, int max_iter, hipStream_t s)
{
    resume(r64_dev, rho_dev, tgt_dev, sel_dev, inner_tol * inner_tol, max_iter, s);
}

void CGEngine::finish_device(const int *sel_dev, double *x, hipStream_t s)
{
    if (precision == FEMD_FP64) finish_t<double>(*this, sel_dev, x, s);
    else finish_t<float>(*this, sel_dev, x, s);
}

// Squared norms of b - A x per system (of b when x is NULL), for the systems with sel; the
// residual stays in r64.
std::vector<double> CGEngine::norms2(const double *b, const double *x, hipStream_t s)
{
    if (x)
    {
        const SpmvPlan &p64 = spmv_plan(*A->pat, FEMD_FP64, A->v.p, s);
        if (A->pat->segments.empty()) spmv_run64(p64, *A->pat, A->v.as<double>(), x, t64.as<double>(), nullptr, s);
        else spmv_run64_masked(p64, *A->pat, A->v.as<double>(), x, t64.as<double>(), sel.as<int>(), s);
    }
    hipLaunchKernelGGL(k_cg_residual, dim3(nch), dim3(BT), 0, s, clo.as<int64_t>(), cpid.as<int>(), sel.as<int>(), b,
                       x ? t64.as<double>() : nullptr, r64.as<double>(), cparts.as<double>());
    hipLaunchKernelGGL(k_cg_sys_sum, dim3(np), dim3(RT), 0, s, cptr.as<int>(), sel.as<int>(), cparts.as<double>(),
                       rho_new.as<double>());
    std::vector<double> h(np, 0.0);
    HIPC(hipMemcpyAsync(h.data(), rho_new.p, 8 * size_t(np), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    return h;
}

void CGEngine::put_int(DevBuf &d, const std::vector<int> &v, hipStream_t s)
{
    HIPC(hipMemcpyAsync(d.p, v.data(), 4 * size_t(np), hipMemcpyHostToDevice, s));
}

void CGEngine::begin(const double *b, double *x, double tol, double inner_tol, int max_updates, bool use_graph, hipStream_t s)
{
    if (!(tol >= 0.0) || !(inner_tol > 0.0 && inner_tol < 1.0) || max_updates < 1)
        fail(FEMD_EINVAL, "inner CG: need tol >= 0, 0 < inner_tol < 1, max_updates >= 1");
    Run &R = run_;
    R = Run{};
    R.b = b;
    R.x = x;
    R.tol = tol;
    R.inner_tol = inner_tol;
    R.max_updates = max_updates;
    R.max_iter = static_cast<int>(std::min<int64_t>(1 << 30, std::max<int64_t>(1000, 10 * max_rows)));
    R.use_graph = use_graph && !timer_on();       // sampled kernel timing needs plain launches
    R.sel.assign(np, 1);
    R.retired.assign(np, 0);
    R.conv.assign(np, 0);
    R.updates.assign(np, 0);
    R.rel.assign(np, 0.0);
    R.hs.assign(np, CGSys{});
    std::vector<double> rho(np, 0.0), tg(np, 0.0);
    put_int(sel, R.sel, s);
    R.nb2 = norms2(b, nullptr, s);
    for (int p = 0; p < np; ++p)
        if (R.nb2[p] == 0.0 && roff[p + 1] > roff[p]) launch_fill(roff[p + 1] - roff[p], 0.0, x + roff[p], s);
    const std::vector<double> rr = norms2(b, x, s);
    for (int p = 0; p < np; ++p)
    {
        const double nb = std::sqrt(R.nb2[p]), rn = std::sqrt(rr[p]);
        R.rel[p] = nb > 0.0 ? rn / nb : 0.0;
        if (nb == 0.0 || R.rel[p] <= tol)
        {
            R.conv[p] = 1;
            R.retired[p] = 1;
            R.sel[p] = 0;
        }
        else
        {
            R.sel[p] = 2;
            rho[p] = rn;
            tg[p] = (tol * nb / rn) * (tol * nb / rn);
            ++R.running;
        }
    }
    std::vector<CGSys> z(np);
    std::memset(z.data(), 0, sizeof(CGSys) * np);
    HIPC(hipMemcpyAsync(st.p, z.data(), sizeof(CGSys) * np, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(rho_new.p, rho.data(), 8 * size_t(np), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(tgt.p, tg.data(), 8 * size_t(np), hipMemcpyHostToDevice, s));
    put_int(retired, R.retired, s);
    put_int(sel, R.sel, s);
    resume(r64.as<double>(), rho_new.as<double>(), tgt.as<double>(), sel.as<int>(), inner_tol * inner_tol, R.max_iter, s);
    build_list(&R.retired, s);
    if (!counts) counts = std::make_unique<Pinned>(2 * sizeof(int));
    HIPC(hipStreamSynchronize(s));                // the host vectors above go out of scope
}

void CGEngine::launch(hipStream_t s)
{
    if (run_.running == 0) return;
    replay(run_.use_graph, s);
    hipLaunchKernelGGL(k_cg_count, dim3(1), dim3(BT), 0, s, np, static_cast<const CGSys *>(st.as<CGSys>()), retired.as<int>(),
                       cnt.as<int>());
    HIPC(hipMemcpyAsync(counts->p, cnt.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    run_.launched = true;
}

std::vector<int> CGEngine::complete(hipStream_t s)
{
    Run &R = run_;
    std::vector<int> done;
    if (!R.launched) return done;
    R.launched = false;
    HIPC(hipStreamSynchronize(s));
    const int running_now = counts->as<int>()[0], paused = counts->as<int>()[1];
    if (paused == 0 && running_now > 0) return done;
    HIPC(hipMemcpyAsync(R.hs.data(), st.p, sizeof(CGSys) * np, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (paused == 0)
    {
        // nothing runs and nothing waits for an update: whatever is left cannot go on
        for (int p = 0; p < np; ++p)
            if (!R.retired[p])
            {
                R.retired[p] = 1;
                done.push_back(p);
            }
        R.running = 0;
        return done;
    }
    // reliable updates of the paused systems
    ++R.rounds;
    for (int p = 0; p < np; ++p) R.sel[p] = (!R.retired[p] && !R.hs[p].active && R.hs[p].reason) ? 1 : 0;
    put_int(sel, R.sel, s);
    finish_device(sel.as<int>(), R.x, s);
    const std::vector<double> rr = norms2(R.b, R.x, s);
    std::vector<double> rho(np, 0.0), tg(np, 0.0);
    for (int p = 0; p < np; ++p)
    {
        if (!R.sel[p]) continue;
        ++R.updates[p];
        const double rn = std::sqrt(rr[p]), nb = std::sqrt(R.nb2[p]);
        R.rel[p] = rn / nb;
        if (R.rel[p] <= R.tol || R.hs[p].total >= R.max_iter || R.updates[p] >= R.max_updates)
        {
            R.conv[p] = R.rel[p] <= R.tol ? 1 : 0;
            R.retired[p] = 1;
            R.sel[p] = 0;
            done.push_back(p);
            continue;
        }
        R.sel[p] = R.hs[p].reason == 2 ? 2 : 1;
        rho[p] = rn;
        tg[p] = (R.tol * nb / rn) * (R.tol * nb / rn);
    }
    HIPC(hipMemcpyAsync(rho_new.p, rho.data(), 8 * size_t(np), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(tgt.p, tg.data(), 8 * size_t(np), hipMemcpyHostToDevice, s));
    put_int(retired, R.retired, s);
    put_int(sel, R.sel, s);
    resume(r64.as<double>(), rho_new.as<double>(), tgt.as<double>(), sel.as<int>(), R.inner_tol * R.inner_tol, R.max_iter, s);
    R.running = 0;
    for (int p = 0; p < np; ++p) R.running += R.retired[p] ? 0 : 1;
    // once a quarter of the listed tiles belong to retired systems, list only the others
    int live = 0;
    for (int p = 0; p < np; ++p)
        if (!R.retired[p]) live += plan->seg_tiles[size_t(p) + 1] - plan->seg_tiles[size_t(p)];
    if (4 * live <= 3 * list_count) build_list(&R.retired, s);
    HIPC(hipStreamSynchronize(s));                // rho, tg go out of scope
    return done;
}

void CGEngine::system_result(int p, int *iters, double *res, int *converged) const
{
    if (iters) *iters = run_.hs[p].total;
    if (res) *res = run_.rel[p];
    if (converged) *converged = run_.conv[p];
}

CGEngine::Result CGEngine::result(int *iters, double *res, hipStream_t s)
{
    HIPC(hipMemcpyAsync(run_.hs.data(), st.p, sizeof(CGSys) * np, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    Result out;
    int nconv = 0;
    for (int p = 0; p < np; ++p)
    {
        if (iters) iters[p] = run_.hs[p].total;
        if (res) res[p] = run_.rel[p];
        nconv += run_.conv[p];
        out.worst = std::max(out.worst, run_.rel[p]);
        out.iterations += run_.hs[p].total;
        out.updates += run_.updates[p];
    }
    out.rounds = run_.rounds;
    out.converged = nconv == np;
    return out;
}

CGEngine::Result CGEngine::solve(const double *b, double *x, double tol, double inner_tol, int max_updates, bool use_graph,
                                 int *iters, double *res, hipStream_t s)
{
    begin(b, x, tol, inner_tol, max_updates, use_graph, s);
    while (!finished())
    {
        launch(s);
        (void)complete(s);
    }
    return result(iters, res, s);
}

} // namespace femd_amd

extern "C" {

int femd_inner_timer(int enable)
{
    using namespace femd_amd;
    return api([&] {
        std::lock_guard<std::mutex> lk(g_timer.mu);
        g_timer.on = enable != 0;
        g_timer.used = 0;
        g_timer.launches = 0;
        g_timer.shape.clear();
    });
}

int femd_inner_timer_read(int64_t *launches, double *total_ms)
{
    using namespace femd_amd;
    return api([&] {
        std::lock_guard<std::mutex> lk(g_timer.mu);
        double t = 0.0;
        int64_t k = 0;
        for (size_t i = 0; i < g_timer.used; ++i)
        {
            const Sample &x = g_timer.samples[i];
            HIPC(hipEventSynchronize(x.e1));
            float ms = 0.0f;
            HIPC(hipEventElapsedTime(&ms, x.e0, x.e1));
            // a launch of a finished solve returns at once: keep the ones that did the product
            if (ms > 0.002f)
            {
                t += ms;
                ++k;
            }
        }
        if (launches) *launches = k;
        if (total_ms) *total_ms = t;
    });
}

int femd_inner_timer_shape(char *name, int len)
{
    using namespace femd_amd;
    return api([&] {
        if (!name || len < 1) fail(FEMD_EINVAL, "femd_inner_timer_shape: need a buffer");
        std::lock_guard<std::mutex> lk(g_timer.mu);
        const size_t k = std::min(g_timer.shape.size(), size_t(len - 1));
        std::memcpy(name, g_timer.shape.data(), k);
        name[k] = '\0';
    });
}

} // extern "C"
