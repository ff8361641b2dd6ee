// sweep.hip -- the triangular sweeps of ILU(0) and SSOR (sparse/precond.hpp:125-144, 189-205) as one
// workgroup's program over a planned stream (sweep.hpp).  The workgroup's twelve waves are four teams
// of three; team s % 4 runs step s.  Each wave holds its team's coming step in registers -- per lane
// its row's input, destination, slot, count, diagonal, the first SW_EREG entry values and the slots
// of their columns -- loaded from HBM during the step after its team's last one, three steps ahead
// of use, so that no data but the z values passes through the LDS: after a barrier the team gathers
// the z values its rows need (one round of LDS reads), subtracts the products in stored order,
// and stores its results.  Rows longer than SW_EREG load their remaining entries when they run.
//
// The steps are separated by a bare s_barrier after the LDS traffic has drained: a __syncthreads
// would also wait for the loads in flight, which are meant to stay in flight across three barriers.

#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace femd_amd {

namespace {

// 16 bytes as the compiler's own vector type (kept in registers, unlike HIP's uint4 struct)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One wave's part of a sweep for its team: after team barriers (the steps of the teams before
// it), the team's steps in groups of SW_TEAMS, straight-line within a group (compute, load the next
// block, wait, wait), so that the registers it loads are the ones it later reads, with no copy
// between them that would wait for a load in flight.
//
// KIND: SW_ILU_FWD (z = r - sum L z), SW_ILU_BWD (z = (y - sum U z) / d), SW_SSOR_FWD (z = (r - sum L z) w / d,
// the row's output scaled by (c d) / w for the backward sweep), SW_SSOR_BWD (z = (y - sum U z) w / d).
// out: where the results go (the backward stream as doubles for a forward sweep, the output vector for a
// backward one).
template <int KIND, int EREG>
__device__ __forceinline__ void sweep_wave(int nsteps, int team, const u32x4 *__restrict__ stream, uint32_t zunit, u32x4 rec,
                                           double w, double c, double *out, const uint8_t *zsb, double *zs, int lane)
{
    constexpr int OREG = EREG / 8;
    // this wave's block of its team's coming step: where it is, its bound and lanes, the record of
    // the block after it, and per lane the row's fields and first entries
    uint32_t base = 0, e = 0, nr = 0, L = 0, pos = 0, sc = 0;
    double vin = 0.0, d = 1.0;
    u32x4 vq[EREG / 2], sq[OREG];
    const uint32_t l16 = 16 * uint32_t(lane);
    // Only the lanes with a row load (the others keep what they had and store nothing), and only
    // the pairs and octets inside the block's bound; the registered places past it are set to
    // zeros.  Each load is a scalar address plus the lane's offset.  The record is read into
    // scalars before the loads replace it.
    auto load_block = [&]() {
        base = uint32_t(__builtin_amdgcn_readfirstlane(int(rec.x)));
        const uint32_t en = uint32_t(__builtin_amdgcn_readfirstlane(int(rec.y)));
        e = en & 0xffffu;
        nr = en >> 16;
        L = uint32_t(__builtin_amdgcn_readfirstlane(int(rec.z)));
        pos = uint32_t(__builtin_amdgcn_readfirstlane(int(rec.w)));
        const uint8_t *b = reinterpret_cast<const uint8_t *>(stream + base);
        // the record through a vector load (an offset the compiler cannot see is zero): a scalar
        // load would be waited for at the barrier, a full memory latency every step
        uint32_t zv;
        __asm__ volatile("v_mov_b32 %0, 0" : "=v"(zv));
        rec = *reinterpret_cast<const u32x4 *>(b + zv);
        const uint32_t np = (e + 1) >> 1, no = (e + 7) >> 3;
        const uint8_t *pairs = b + 16 * sw_u_pairs(KIND, int(L)), *octs = pairs + 16 * np * L;
        if (uint32_t(lane) < nr)
        {
            // the row's input and slot | count: one unit per lane
            const u32x4 rw = *reinterpret_cast<const u32x4 *>(b + sw_b_vin(0) + l16);
            vin = __hiloint2double(int(rw.y), int(rw.x));
            sc = rw.w;
            if (sweep_has_diag(KIND)) d = *reinterpret_cast<const double *>(b + sw_b_diag(int(L), 0) + 8 * uint32_t(lane));
#pragma unroll
            for (int q = 0; q < EREG / 2; ++q)
            {
                if (uint32_t(q) < np) vq[q] = *reinterpret_cast<const u32x4 *>(pairs + 16 * uint32_t(q) * L + l16);
                else vq[q] = u32x4{0u, 0u, 0u, 0u};
            }
#pragma unroll
            for (int r = 0; r < OREG; ++r)
            {
                if (uint32_t(r) < no) sq[r] = *reinterpret_cast<const u32x4 *>(octs + 16 * uint32_t(r) * L + l16);
                else sq[r] = u32x4{0u, 0u, 0u, 0u};
            }
        }
    };
    // The same instructions for a wave without rows (its registers hold zeros then, its lanes
    // store nothing): a branch around them would leave the compiler unsure whether the loads are
    // done, and it would then wait for the next block's loads in the middle of issuing them.
    auto compute = [&]() {
        // the loaded registers are first touched here: without this the compiler moves work on
        // them (sc & 0xffff, say) up into the load step, where it waits for the loads it issued
        __asm__ volatile("" : "+v"(sc), "+v"(vin), "+v"(d));
#pragma unroll
        for (int r = 0; r < OREG; ++r) __asm__ volatile("" : "+v"(sq[r]));
#pragma unroll
        for (int q = 0; q < EREG / 2; ++q) __asm__ volatile("" : "+v"(vq[q]));
        double acc = vin;
        double zz[EREG];
        // every gather of the registered entries at once: the octets hold the z values' LDS byte
        // offsets (a padded entry's is slot SW_ZERO's, +0.0)
#pragma unroll
        for (int k = 0; k < EREG; ++k)
        {
            const uint32_t word = sq[k >> 3][(k >> 1) & 3];
            const uint32_t off = (k & 1) ? word >> 16 : word & 0xffffu;
            zz[k] = *reinterpret_cast<const double *>(zsb + off);
        }
        // keep the gathers together, ahead of the products: one round trip for all of them
        __builtin_amdgcn_sched_barrier(0);
        // the products in stored order: the reference's s -= v[p] * z[ix[p]], one rounding for the
        // product and one for the difference; a padded entry subtracts +0.0, which leaves acc as it is
#pragma unroll
        for (int k = 0; k < EREG; ++k)
        {
            const u32x4 pq = vq[k >> 1];
            const uint32_t lo = (k & 1) ? pq.z : pq.x, hi = (k & 1) ? pq.w : pq.y;
            const double prod = __hiloint2double(int(hi), int(lo)) * zz[k];
            acc = acc - prod;
        }
        // entries past EREG (rare): loaded now, padded like the first ones up to the block's bound
        if (e > uint32_t(EREG))
        {
            const uint32_t np = (e + 1) >> 1;
            const uint8_t *pairs = reinterpret_cast<const uint8_t *>(stream + base) + 16 * sw_u_pairs(KIND, int(L)),
                          *octs = pairs + 16 * np * L;
            for (int k0 = EREG; k0 < int(e); k0 += 8)
            {
                const u32x4 so = *reinterpret_cast<const u32x4 *>(octs + 16 * uint32_t(k0 >> 3) * L + l16);
                double zk[8], vk[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                {
                    const uint32_t off = (u & 1) ? so[u >> 1] >> 16 : so[u >> 1] & 0xffffu;
                    zk[u] = *reinterpret_cast<const double *>(zsb + off);
                    const uint32_t q = min(uint32_t(k0 + u) >> 1, np - 1);
                    vk[u] = *reinterpret_cast<const double *>(pairs + 16 * q * L + l16 + 8 * (u & 1));
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                {
                    const double prod = vk[u] * zk[u];
                    acc = k0 + u < int(e) ? acc - prod : acc;
                }
            }
        }
        if (uint32_t(lane) < nr)
        {
            double z = acc, res = acc;
            if (KIND == SW_ILU_BWD) z = res = acc / d;
            else if (KIND == SW_SSOR_FWD || KIND == SW_SSOR_BWD)
            {
                z = acc * w / d;
                res = KIND == SW_SSOR_FWD ? z * (c * d / w) : z;
            }
            zs[sc & 0xffffu] = z;
            // the results in the sweep's order: 64 consecutive doubles per wave
            out[pos + uint32_t(lane)] = res;
        }
    };
    // LDS settled, global loads and stores left in flight
    auto barrier = []() { __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); };

    load_block();
    for (int s = 0; s < team && s < nsteps; ++s) barrier();
    int s = team;
    for (; s + SW_TEAMS <= nsteps; s += SW_TEAMS)
    {
        compute();
        barrier();
        load_block();
#pragma unroll
        for (int k = 1; k < SW_TEAMS; ++k) barrier();
    }
    if (s < nsteps)
    {
        compute();
        barrier();
        for (++s; s < nsteps; ++s) barrier();
    }
}

// first: the teams' first blocks (a record per wave); zunit: the stream's zeros; EREG: entries per
// row in registers
template <int KIND, int EREG>
__global__ __launch_bounds__(SW_THREADS) void k_sweep(int nsteps, const u32x4 *__restrict__ stream,
                                                      const u32x4 *__restrict__ first, uint32_t zunit, double w, double c,
                                                      double *out, const int *stop)
{
    __shared__ double zs[SW_SLOTS];
    if (stop && *stop) return;
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
    if (t == 0) zs[SW_ZERO] = 0.0;
    const u32x4 rec = first[wv];
    __syncthreads();
    // the team as a scalar: the steps' loop runs on scalar counters, not per lane
    sweep_wave<KIND, EREG>(nsteps, __builtin_amdgcn_readfirstlane(wv / SW_TWAVES), stream, zunit, rec, w, c, out,
                     reinterpret_cast<const uint8_t *>(zs), zs, lane);
}

// dst[to[k]] = src[k]: a sweep's results (in its order) to where they go -- the backward stream's
// vin fields, or the output vector through the ordering
__global__ void k_sweep_scatter(int n, const int *__restrict__ to, const double *__restrict__ src, double *dst, const int *stop)
{
    if (stop && *stop) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[to[k]] = src[k];
}

// stream[in_pos[k]] = src[perm ? perm[k] : k]: the rows' right-hand side into the forward stream
__global__ void k_sweep_fill(int n, const int *__restrict__ in_pos, const int *__restrict__ perm,
                             const double *__restrict__ src, double *stream, const int *stop)
{
    if (stop && *stop) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) stream[in_pos[k]] = src[perm ? perm[k] : k];
}

// the factor's (or the matrix's) values into the streams: fwd[fmap[p]], bwd[bmap[p]] = v[p]
__global__ void k_sweep_pack(int64_t nnz, const int *__restrict__ fmap, const int *__restrict__ bmap,
                             const double *__restrict__ v, double *fwd, double *bwd)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    if (fmap[p] >= 0) fwd[fmap[p]] = v[p];
    if (bmap[p] >= 0) bwd[bmap[p]] = v[p];
}

} // namespace

template <int KIND>
void launch_kind(int ereg, int nsteps, const u32x4 *st, const u32x4 *fr, uint32_t zunit, double omega, double c, double *out,
                 const int *stop, hipStream_t s)
{
    if (ereg <= 8) k_sweep<KIND, 8><<<1, SW_THREADS, 0, s>>>(nsteps, st, fr, zunit, omega, c, out, stop);
    else if (ereg <= 16) k_sweep<KIND, 16><<<1, SW_THREADS, 0, s>>>(nsteps, st, fr, zunit, omega, c, out, stop);
    else k_sweep<KIND, 24><<<1, SW_THREADS, 0, s>>>(nsteps, st, fr, zunit, omega, c, out, stop);
}

void launch_sweep(int kind, int ereg, int nsteps, const void *stream, const void *first, uint32_t zunit, double omega,
                  double *out, const int *stop, hipStream_t s)
{
    if (nsteps <= 0) return;
    const double c = (2.0 - omega) / omega;
    const u32x4 *st = static_cast<const u32x4 *>(stream), *fr = static_cast<const u32x4 *>(first);
    switch (kind)
    {
    case SW_ILU_FWD: launch_kind<SW_ILU_FWD>(ereg, nsteps, st, fr, zunit, omega, c, out, stop, s); break;
    case SW_ILU_BWD: launch_kind<SW_ILU_BWD>(ereg, nsteps, st, fr, zunit, omega, c, out, stop, s); break;
    case SW_SSOR_FWD: launch_kind<SW_SSOR_FWD>(ereg, nsteps, st, fr, zunit, omega, c, out, stop, s); break;
    default: launch_kind<SW_SSOR_BWD>(ereg, nsteps, st, fr, zunit, omega, c, out, stop, s); break;
    }
}

void launch_sweep_fill(int n, const int *in_pos, const int *perm, const double *src, double *stream, const int *stop,
                       hipStream_t s)
{
    if (n > 0) k_sweep_fill<<<(n + 255) / 256, 256, 0, s>>>(n, in_pos, perm, src, stream, stop);
}

void launch_sweep_scatter(int n, const int *to, const double *src, double *dst, const int *stop, hipStream_t s)
{
    if (n > 0) k_sweep_scatter<<<(n + 255) / 256, 256, 0, s>>>(n, to, src, dst, stop);
}

void launch_sweep_pack(int64_t nnz, const int *fmap, const int *bmap, const double *v, double *fwd, double *bwd, hipStream_t s)
{
    if (nnz > 0) k_sweep_pack<<<unsigned((nnz + 255) / 256), 256, 0, s>>>(nnz, fmap, bmap, v, fwd, bwd);
}

} // namespace femd_amd
