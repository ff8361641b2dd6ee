// sweep.hip -- the triangular sweeps of ILU(0) and SSOR (sparse/precond.hpp:125-144, 189-205) as one
// workgroup's program over a planned stream (sweep.hpp): eight consumer waves compute the rows of a
// step and store their results; four producer waves keep an LDS ring filled from HBM, each
// carrying the blocks of every fourth step in registers for four steps.
//
// A consumer wave holds the next step's rows in registers before the step begins -- each lane its
// row's input, destination, slot, count, diagonal, up to SW_EREG values and their slots, read from
// the ring during the step before -- so that after a barrier a step costs one round of LDS gathers
// of z values, the products subtracted in stored order, and the stores of its results; the reads
// of the step after it are issued while those products run.  Rows longer than SW_EREG read their
// remaining entries from the ring during their own step.
//
// The steps are separated by a bare s_barrier after the LDS traffic has drained: a __syncthreads
// would also wait for the producers' loads in flight, which are meant to stay in flight across
// four barriers.  Consumer waves issue no global loads and producer waves no global stores, so a
// producer's wait before its ring stores covers only its own loads of four steps before, and the
// consumers' stores never hold anybody up.

#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace femd_amd {

namespace {

// 16 bytes as the compiler's own vector type (kept in registers, unlike HIP's uint4 struct)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// KIND: SW_ILU_FWD (z = r - sum L z), SW_ILU_BWD (z = (y - sum U z) / d), SW_SSOR_FWD (z = (r - sum L z) w / d,
// the row's output scaled by (c d) / w for the backward sweep), SW_SSOR_BWD (z = (y - sum U z) w / d).
// out: where the results go (the backward stream as doubles for a forward sweep, the output vector for a
// backward one).  The LDS holds the z slots (zbytes) and then the ring (rbytes, whole blocks); ring
// offsets are kept below rbytes, and an offset at most one ring past it is brought back by wrap().
template <int KIND>
__global__ __launch_bounds__(SW_THREADS) void k_sweep(int nsteps, const u32x4 *__restrict__ stream, const int *__restrict__ wb,
                                                      uint32_t zbytes, uint32_t rbytes, double w, double c, double *out,
                                                      const int *stop)
{
    __shared__ u32x4 lds4[SW_LDS / 16];
    if (stop && *stop) return;
    double *zs = reinterpret_cast<double *>(lds4);
    uint8_t *ring = reinterpret_cast<uint8_t *>(lds4) + zbytes;
    const uint32_t R = rbytes, NB = rbytes / SW_BLOCK;
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
    const bool producer = wv >= SW_CWAVES;
    const int p = wv - SW_CWAVES;
    auto wrap = [R](uint32_t at) { return min(at, at - R); };
    auto f64 = [&](uint32_t at) { return *reinterpret_cast<const double *>(ring + wrap(at)); };
    auto u32 = [&](uint32_t at) { return *reinterpret_cast<const uint32_t *>(ring + wrap(at)); };
    auto v4 = [&](uint32_t at) { return *reinterpret_cast<const u32x4 *>(ring + wrap(at)); };

    if (t == 0) zs[SW_ZERO] = 0.0;
    const int b0 = wb[0];
    for (int b = wv; b < b0; b += SW_CWAVES + SW_PWAVES)
        reinterpret_cast<u32x4 *>(ring + uint32_t(b) * SW_BLOCK)[lane] = stream[int64_t(b) * (SW_BLOCK / 16) + lane];
    __syncthreads();

    if (producer)
    {
        // A turn's loads are straight-line code, every one issued (past the turn's blocks it loads
        // its last one again; the stream ends with a spare block, so a turn with none loads that):
        // a load under a branch of its own makes the compiler wait for the loads before it.
        u32x4 buf[SW_PMAX];
        auto load_turn = [&](int first, int count) {
            const int last = count > 0 ? count - 1 : 0;
            const u32x4 *src = stream + int64_t(first) * (SW_BLOCK / 16) + lane;
#pragma unroll
            for (int k = 0; k < SW_PMAX; ++k) buf[k] = src[int64_t(k < last ? k : last) * (SW_BLOCK / 16)];
        };
        int bfirst = wb[p], nbuf = wb[p + 1] - wb[p];
        load_turn(bfirst, nbuf);
        uint32_t o = 0;
        for (int s = 0; s < nsteps; ++s)
        {
            if ((s & (SW_PWAVES - 1)) == p)
            {
                uint32_t at = (uint32_t(bfirst) % NB) * SW_BLOCK;
#pragma unroll
                for (int k = 0; k < SW_PMAX; ++k)
                {
                    if (k < nbuf) reinterpret_cast<u32x4 *>(ring + at)[lane] = buf[k];
                    at = wrap(at + SW_BLOCK);
                }
                // the blocks of step s + 4, named by this step's header
                bfirst = int(u32(o + 8));
                nbuf = int(u32(o + 12)) - bfirst;
                load_turn(bfirst, nbuf);
            }
            o = wrap(o + u32(o));
            // LDS settled, global loads left in flight
            __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        }
        return;
    }

    // The consumer's registers for a step: the step's ring offset, size and waves; this wave's
    // block, entry bound and lanes; per lane the row's input, destination, slot | count << 16,
    // diagonal, the pre table (lane q holds pre[q]), the first SW_EREG values (pairs) and their
    // slots (octets).  Producers and consumers run loops of their own, one barrier per step in
    // each, so that the registers one role carries across steps are not held by the other.
    uint32_t o = 0, size = 0, nw = 0, blk = 0, e = 0, L = 0, sc = 0, pv = 0;
    int dst = -1;
    double vin = 0.0, d = 1.0;
    u32x4 vq[SW_EREG / 2], sq[SW_OREG];
    // a step's header, then this wave's rows, then its first entries: three rounds of LDS reads
    auto read_header = [&](uint32_t at, uint32_t &sz, uint32_t &nwv, u32x4 &rec) {
        sz = u32(at);
        nwv = u32(at + 4);
        rec = v4(at + 16 + 16 * wv);
    };
    auto read_rows = [&](uint32_t at, const u32x4 &rec) {
        blk = wrap(at + rec.x);
        e = rec.y;
        L = rec.z;
        const uint32_t lq = uint32_t(lane) < L ? uint32_t(lane) : 0u;
        vin = f64(blk + 8 * lq);
        dst = int(u32(blk + uint32_t(sw_off_dest(int(L))) + 4 * lq));
        sc = u32(blk + uint32_t(sw_off_sc(int(L))) + 4 * lq);
        if (sweep_has_diag(KIND)) d = f64(blk + uint32_t(sw_off_diag(int(L))) + 8 * lq);
        pv = u32(blk + uint32_t(sw_off_pre(KIND, int(L))) + 4 * (uint32_t(lane) < rec.w ? uint32_t(lane) : 0u));
    };
    auto pair_at = [&](int q) { return blk + 16 * (uint32_t(__builtin_amdgcn_readlane(int(pv), q)) + uint32_t(lane)); };
    auto oct_at = [&](int r) {
        const int np2 = int(e + 1) >> 1;
        return blk + 16 * (uint32_t(__builtin_amdgcn_readlane(int(pv), np2 + r)) + uint32_t(lane));
    };
    // the first SW_EREG entries of this wave's rows (every read issued: past a row's end, or past
    // the wave's block, a lane reads bytes it never uses, all of them in the ring)
    auto read_entries = [&]() {
#pragma unroll
        for (int q = 0; q < SW_EREG / 2; ++q) vq[q] = v4(pair_at(q));
#pragma unroll
        for (int r = 0; r < SW_OREG; ++r) sq[r] = v4(oct_at(r));
    };
    {
        u32x4 rec;
        read_header(0, size, nw, rec);
        if (wv < int(nw))
        {
            read_rows(0, rec);
            read_entries();
        }
    }
    for (int s = 0; s < nsteps; ++s)
    {
        // the next step's header now (it is in the ring); its rows once this step's are done
        const uint32_t o2 = wrap(o + size);
        uint32_t size2, nw2;
        u32x4 rec2;
        read_header(o2, size2, nw2, rec2);
        if (wv < int(nw))
        {
            const int cnt = int(sc >> 16);
            double acc = vin;
            double zz[SW_EREG];
            // every gather of the registered entries at once (past its row's end a lane reads slot
            // SW_ZERO, whose 0.0 it never uses)
#pragma unroll
            for (int k = 0; k < SW_EREG; ++k)
            {
                const uint32_t raw = (sq[k >> 3][(k >> 1) & 3] >> (16 * (k & 1))) & (SW_SLOTS - 1);
                zz[k] = zs[k < cnt ? raw : uint32_t(SW_ZERO)];
            }
            // keep the gathers together, ahead of the products: one round trip for all of them
            __builtin_amdgcn_sched_barrier(0);
            // the products in stored order: the reference's s -= v[p] * z[ix[p]], one rounding for
            // the product and one for the difference; lanes past their row's end keep acc
#pragma unroll
            for (int k = 0; k < SW_EREG; ++k)
            {
                const u32x4 pq = vq[k >> 1];
                const uint32_t lo = (k & 1) ? pq.z : pq.x, hi = (k & 1) ? pq.w : pq.y;
                const double prod = __hiloint2double(int(hi), int(lo)) * zz[k];
                acc = k < cnt ? acc - prod : acc;
            }
            // entries past SW_EREG (rare): straight from the ring, this step's bytes
            if (int(e) > SW_EREG)
            {
                for (int k0 = SW_EREG; k0 < int(e); k0 += 8)
                {
                    const u32x4 so = v4(oct_at(k0 >> 3));
                    double zk[8], vk[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const uint32_t raw = (so[u >> 1] >> (16 * (u & 1))) & (SW_SLOTS - 1);
                        zk[u] = zs[k0 + u < cnt ? raw : uint32_t(SW_ZERO)];
                        vk[u] = f64(pair_at((k0 + u) >> 1) + 8 * (u & 1));
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const double prod = vk[u] * zk[u];
                        acc = k0 + u < cnt ? acc - prod : acc;
                    }
                }
            }
            if (uint32_t(lane) < L)
            {
                double z = acc, res = acc;
                if (KIND == SW_ILU_BWD) z = res = acc / d;
                else if (KIND == SW_SSOR_FWD || KIND == SW_SSOR_BWD)
                {
                    z = acc * w / d;
                    res = KIND == SW_SSOR_FWD ? z * (c * d / w) : z;
                }
                zs[sc & 0xffffu] = z;
                if (dst >= 0) out[dst] = res;
            }
        }
        // the next step's rows and first entries, into the registers this step is done with
        o = o2;
        size = size2;
        nw = nw2;
        if (wv < int(nw))
        {
            read_rows(o, rec2);
            read_entries();
        }
        // LDS settled, global stores left in flight
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

void launch_sweep(int kind, int nsteps, const void *stream, const int *wb, int zbytes, int rbytes, double omega, double *out,
                  const int *stop, hipStream_t s)
{
    if (nsteps <= 0) return;
    const double c = (2.0 - omega) / omega;
    const u32x4 *st = static_cast<const u32x4 *>(stream);
    const uint32_t zb = uint32_t(zbytes), rb = uint32_t(rbytes);
    switch (kind)
    {
    case SW_ILU_FWD: k_sweep<SW_ILU_FWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, zb, rb, omega, c, out, stop); break;
    case SW_ILU_BWD: k_sweep<SW_ILU_BWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, zb, rb, omega, c, out, stop); break;
    case SW_SSOR_FWD: k_sweep<SW_SSOR_FWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, zb, rb, omega, c, out, stop); break;
    default: k_sweep<SW_SSOR_BWD><<<1, SW_THREADS, 0, s>>>(nsteps, st, wb, zb, rb, omega, c, out, stop); break;
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
