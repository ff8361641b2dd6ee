// sweep.hip -- the triangular sweeps of ILU(0) and SSOR (sparse/precond.hpp:125-144, 189-205) as one
// workgroup's program over a planned stream (sweep.hpp): eight consumer waves compute the rows of a
// step from LDS (z slots, the step's bytes in the ring) and store their results; four producer
// waves keep the ring filled from HBM, each carrying the blocks of every fourth step in registers
// for four steps.
//
// The steps are separated by a bare s_barrier after the LDS traffic has drained: a __syncthreads
// would also wait for the producers' loads in flight, which are meant to stay in flight across
// four barriers.  Consumer waves issue no global loads and producer waves no global stores, so a
// producer's wait before its ring stores covers only its own loads of four steps before, and the
// consumers' stores never hold anybody up.  A step lasts as long as its longest row: each lane
// walks its row eight entries at a time, every ring read unconditional (each ring address is in
// the ring; an entry past the row's end reads whatever is there and is not used), so the wave
// takes no branch per entry.
#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace femd_amd {

namespace {

constexpr uint32_t RING_MASK = SW_RING - 1;
// 16 bytes as the compiler's own vector type (kept in registers, unlike HIP's uint4 struct)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline double lds_f64(const uint8_t *ring, uint32_t at)
{
    return *reinterpret_cast<const double *>(ring + (at & RING_MASK));
}
__device__ inline uint32_t lds_u32(const uint8_t *ring, uint32_t at)
{
    return *reinterpret_cast<const uint32_t *>(ring + (at & RING_MASK));
}
__device__ inline int32_t lds_i32(const uint8_t *ring, uint32_t at)
{
    return *reinterpret_cast<const int32_t *>(ring + (at & RING_MASK));
}
__device__ inline uint32_t lds_u16(const uint8_t *ring, uint32_t at)
{
    return *reinterpret_cast<const uint16_t *>(ring + (at & RING_MASK));
}

// KIND: SW_ILU_FWD (z = r - sum L z), SW_ILU_BWD (z = (y - sum U z) / d), SW_SSOR_FWD (z = (r - sum L z) w / d,
// the row's output scaled by (c d) / w for the backward sweep), SW_SSOR_BWD (z = (y - sum U z) w / d).
// out: where the results go (the backward stream as doubles for a forward sweep, the output vector for a
// backward one).
template <int KIND>
__global__ __launch_bounds__(SW_THREADS) void k_sweep(int nsteps, const u32x4 *__restrict__ stream, const int *__restrict__ wb,
                                                      double w, double c, double *out, const int *stop)
{
    __shared__ u32x4 ring4[SW_RING / 16];
    __shared__ double zs[SW_SLOTS];
    if (stop && *stop) return;
    const uint8_t *ring = reinterpret_cast<const uint8_t *>(ring4);
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
    const bool producer = wv >= SW_CWAVES;
    const int p = wv - SW_CWAVES;

    if (t == 0) zs[SW_ZERO] = 0.0;
    const int b0 = wb[0];
    for (int b = wv; b < b0; b += SW_CWAVES + SW_PWAVES)
        ring4[((uint32_t(b) * SW_BLOCK) & RING_MASK) / 16 + lane] = stream[int64_t(b) * (SW_BLOCK / 16) + lane];
    u32x4 buf[SW_PMAX];
    int bfirst = 0, nbuf = 0;
    // A turn's loads are straight-line code, every one issued (past the turn's blocks it loads its
    // last one again; the stream ends with a spare block, so a turn with none loads that): a load
    // under a branch of its own makes the compiler wait for the loads before it.
    auto load_turn = [&](int first, int count) {
        const int last = count > 0 ? count - 1 : 0;
        const u32x4 *src = stream + int64_t(first) * (SW_BLOCK / 16) + lane;
#pragma unroll
        for (int k = 0; k < SW_PMAX; ++k) buf[k] = src[int64_t(k < last ? k : last) * (SW_BLOCK / 16)];
    };
    if (producer)
    {
        bfirst = wb[p];
        nbuf = wb[p + 1] - bfirst;
        load_turn(bfirst, nbuf);
    }
    __syncthreads();

    // the consumers' view of the current step: its offset, size and waves, and this wave's record
    uint32_t o = 0, size = 0, nw = 0;
    u32x4 rec = {0, 0, 0, 0};
    if (!producer)
    {
        size = lds_u32(ring, 0);
        nw = lds_u32(ring, 4);
        rec = *reinterpret_cast<const u32x4 *>(ring + 16 + 16 * wv);
    }
    for (int s = 0; s < nsteps; ++s)
    {
        if (producer)
        {
            if ((s & (SW_PWAVES - 1)) == p)
            {
#pragma unroll
                for (int k = 0; k < SW_PMAX; ++k)
                    if (k < nbuf) ring4[((uint32_t(bfirst + k) * SW_BLOCK) & RING_MASK) / 16 + lane] = buf[k];
                // the blocks of step s + 8, named by this step's header
                bfirst = int(lds_u32(ring, o + 8));
                nbuf = int(lds_u32(ring, o + 12)) - bfirst;
                load_turn(bfirst, nbuf);
            }
            o += lds_u32(ring, o);
        }
        else
        {
            // the next step's header is already in the ring: read it now, use it next step
            const uint32_t o2 = o + size;
            const uint32_t size2 = lds_u32(ring, o2), nw2 = lds_u32(ring, o2 + 4);
            const u32x4 rec2 = *reinterpret_cast<const u32x4 *>(ring + ((o2 + 16 + 16 * wv) & RING_MASK));
            const int L = int(rec.z);
            if (wv < int(nw) && lane < L)
            {
                const int e = int(rec.y);
                const uint32_t b = o + rec.x;
                const uint32_t po = b + uint32_t(sw_off_pre(KIND, L));
                const uint32_t eo = b + uint32_t(sw_off_ev(KIND, L, e)), so = b + uint32_t(sw_off_es(KIND, L, e, rec.w));
                double acc = lds_f64(ring, b + 8 * lane);
                const int dst = lds_i32(ring, b + uint32_t(sw_off_dest(L)) + 4 * lane);
                const uint32_t own = lds_u16(ring, b + uint32_t(sw_off_slot(L)) + 2 * lane);
                const int cnt = int(lds_u16(ring, b + uint32_t(sw_off_cnt(L)) + 2 * lane));
                double d = 1.0;
                if (sweep_has_diag(KIND)) d = lds_f64(ring, b + uint32_t(sw_off_diag(L)) + 8 * lane);
                // the row's products eight at a time (their loads in flight together), subtracted in
                // stored order: the reference's s -= v[p] * z[ix[p]], one rounding for the product
                // and one for the difference; lanes past their row's end keep acc
                for (int k0 = 0; k0 < e; k0 += 8)
                {
                    const u32x4 pa = *reinterpret_cast<const u32x4 *>(ring + ((po + 4 * k0) & RING_MASK));
                    const u32x4 pb = *reinterpret_cast<const u32x4 *>(ring + ((po + 4 * k0 + 16) & RING_MASK));
                    double v[8], zz[8];
                    uint32_t sl[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const uint32_t at = (u < 4 ? pa[u & 3] : pb[u & 3]) + uint32_t(lane);
                        v[u] = lds_f64(ring, eo + 8 * at);
                        const uint32_t raw = lds_u16(ring, so + 2 * at);
                        sl[u] = k0 + u < cnt ? raw : uint32_t(SW_ZERO);
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) zz[u] = zs[sl[u] & (SW_SLOTS - 1)];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const double prod = v[u] * zz[u];
                        acc = k0 + u < cnt ? acc - prod : acc;
                    }
                }
                double z = acc, res = acc;
                if (KIND == SW_ILU_BWD) z = res = acc / d;
                else if (KIND == SW_SSOR_FWD || KIND == SW_SSOR_BWD)
                {
                    z = acc * w / d;
                    res = KIND == SW_SSOR_FWD ? z * (c * d / w) : z;
                }
                zs[own] = z;
                if (dst >= 0) out[dst] = res;
            }
            o = o2;
            size = size2;
            nw = nw2;
            rec = rec2;
        }
        // LDS settled, global loads and stores left in flight
        __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
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

void launch_sweep(int kind, int nsteps, const void *stream, const int *wb, double omega, double *out, const int *stop,
                  hipStream_t s)
{
    if (nsteps <= 0) return;
    const double c = (2.0 - omega) / omega;
    const u32x4 *st = static_cast<const u32x4 *>(stream);
    switch (kind)
    {
    case SW_ILU_FWD: k_sweep<SW_ILU_FWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, omega, c, out, stop); break;
    case SW_ILU_BWD: k_sweep<SW_ILU_BWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, omega, c, out, stop); break;
    case SW_SSOR_FWD: k_sweep<SW_SSOR_FWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, omega, c, out, stop); break;
    default: k_sweep<SW_SSOR_BWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, omega, c, out, stop); break;
    }
}

void launch_sweep_fill(int n, const int *in_pos, const int *perm, const double *src, double *stream, const int *stop,
                       hipStream_t s)
{
    if (n > 0) k_sweep_fill<<<(n + 255) / 256, 256, 0, s>>>(n, in_pos, perm, src, stream, stop);
}

void launch_sweep_pack(int64_t nnz, const int *fmap, const int *bmap, const double *v, double *fwd, double *bwd, hipStream_t s)
{
    if (nnz > 0) k_sweep_pack<<<unsigned((nnz + 255) / 256), 256, 0, s>>>(nnz, fmap, bmap, v, fwd, bwd);
}

} // namespace femd_amd
