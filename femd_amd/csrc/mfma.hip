// mfma.hip -- element matrices of a dx group on the matrix cores (FP64 MFMA): the fast assembly
// mode, beside the exact VALU kernel of kernels.hip (k_elem_matrix), whose operation order is the
// reference's.
//
// The contraction of forms/assembly_2d.hpp:102-126 is
//     Ke[i][j] = sum_t sum_q s_t(q) P_{a_t}(i, q) P_{b_t}(j, q),   s_t(q) = w_q |det J| c_t(q),
// with the physical rows P_m = sum_alpha M_{m,alpha} Rhat_alpha: Rhat_0 the reference values,
// Rhat_1 / Rhat_2 the xi / eta derivatives at the point, and M the rows of J^{-T}
// (M_0 = (1, 0, 0), M_1 = (0, i0, i2), M_2 = (0, i1, i3)).  Taking the cell out of the sum:
//     Ke[ij]    = sum_k D_e[k] R[k][ij],         k = (combo c, point q),
//     D_e[c, q] = sum_t s_t(q) M_{a_t, alpha_c} M_{b_t, beta_c},
//     R[c, q][ij] = Rhat_{alpha_c}(i, q) Rhat_{beta_c}(j, q)
// -- one GEMM over the cells (M = cells, K = combos x points, N = local entries) whose right
// operand R is the same for every cell.  When the terms are closed under swapping the two codes
// (with the same coefficient), D is symmetric in (alpha, beta): (alpha, beta) and (beta, alpha)
// share one k with their R tables added, and only the entries i <= j are computed, each written
// to (i, j) and (j, i).
//
// One wave computes 16 cells x 16 local entries per v_mfma_f64_16x16x4f64 (A: the cells' D, four
// k at a time; B: four rows of R).  Lane l holds A[cell l%16][k l/16] and B[k l/16][entry l%16];
// the result's element v of lane l is row (cell) 4 v + l/16, column (entry) l%16.  Every combo's
// points are padded to a multiple of four (spc steps, zero rows of R), so a step belongs to one
// combo: the wave loads its cells' weights and coefficient values at its points once (all loads in
// flight together), forms each combo's physical-row weights once, and D is a few multiply-adds
// per step ahead of the matrix instructions.  R is staged in LDS when it fits in 64 KiB.
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <cmath>

namespace femd_amd {

namespace {
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int MFMA_TPB = 256;

// The Jacobian's inverse and |det| at (xi, eta): constant on a triangle (g = x0, y0, J), the
// bilinear map's on a quadrilateral (g = its four vertices).
struct Jinv {
    double i0, i1, i2, i3, ad;
};
template <int NV>
__device__ inline Jinv jinv(const double *g, double xi, double eta)
{
    double J0, J1, J2, J3;
    if constexpr (NV == 4)
    {
        J0 = (g[2] - g[0]) * (1 - eta) + (g[4] - g[6]) * eta;
        J1 = (g[6] - g[0]) * (1 - xi) + (g[4] - g[2]) * xi;
        J2 = (g[3] - g[1]) * (1 - eta) + (g[5] - g[7]) * eta;
        J3 = (g[7] - g[1]) * (1 - xi) + (g[5] - g[3]) * xi;
    }
    else
    {
        J0 = g[2]; J1 = g[3]; J2 = g[4]; J3 = g[5];
    }
    const double d = J0 * J3 - J1 * J2;
    return Jinv{J3 / d, -J1 / d, -J2 / d, J0 / d, fabs(d)};
}

// M_{m, alpha}: the weight of reference row alpha in physical row m.
__device__ inline double mrow(int m, int alpha, const Jinv &j)
{
    if (m == 0) return alpha == 0 ? 1.0 : 0.0;
    if (alpha == 0) return 0.0;
    if (m == 1) return alpha == 1 ? j.i0 : j.i2;
    return alpha == 1 ? j.i1 : j.i3;
}

// The physical-row weights of combo (al, be) at one point: W0 gathers the constant terms (their
// constants folded in), W[u] the terms whose coefficient is array u.
template <int UMAX>
__device__ inline double combo_weights(const MfmaTerms &T, int al, int be, const Jinv &jc, double *W)
{
    double W0 = 0.0;
#pragma unroll
    for (int u = 0; u < UMAX; ++u) W[u] = 0.0;
    // unrolled over the most terms a group has, so the term table is read into scalar registers
    // once, ahead of every loop
#pragma unroll
    for (int t = 0; t < MAX_TERMS; ++t)
    {
        if (t >= T.nt) break;
        const double m = mrow(T.a[t], al, jc) * mrow(T.b[t], be, jc);
        const int sl = T.slot[t];
        if (sl < 0) W0 += T.cval[t] * m;
#pragma unroll
        for (int u = 0; u < UMAX; ++u)
            if (sl == u) W[u] += m;
    }
    return W0;
}

// Three waves per SIMD (at most 168 registers; two on quadrilaterals or with more than two
// coefficient arrays, which need more): the loads of one wave's group hide behind the matrix work
// of the others.
template <int NV, int NT, int UMAX, bool RLDS>
__global__ __launch_bounds__(MFMA_TPB) __attribute__((amdgpu_waves_per_eu(NV == 4 || UMAX > 2 ? 2 : 3))) void k_elem_matrix_mfma(MfmaJob J, MfmaTerms T)
{
    extern __shared__ double sh[];
    const int npad = 16 * J.ntile;
    const double *R = J.rt;
    if constexpr (RLDS)
    {
        const int total = 4 * J.ks * npad;             // even: npad is a multiple of 16
        const double2 *src = reinterpret_cast<const double2 *>(J.rt);
        for (int k = threadIdx.x; k < total / 2; k += MFMA_TPB) reinterpret_cast<double2 *>(sh)[k] = src[k];
        __syncthreads();
        R = sh;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ci = lane & 15, kk = lane >> 4;
    const int ngroups = (J.nc + 15) >> 4;
    const int64_t N2 = int64_t(J.N) * J.N;
    constexpr int NG = NV == 4 ? 8 : 6;
    // the output map of a single pass of NT tiles, read once
    int2 oc1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
    {
        const int n = t * 16 + ci;
        oc1[t] = n < J.nout ? J.ocol[n] : int2{-1, -1};
    }
    // a group's loads: the cell geometry, and at this lane's points q = 4 j + kk (j < spc) the
    // weights and coefficient values
    struct Loads {
        double gl[NG], wq[MFMA_SPCMAX], cv[UMAX][MFMA_SPCMAX];
    };
    auto load = [&](int g, Loads &L) {
        const int e = g * 16 + ci;
        const bool valid = g < ngroups && e < J.nc;
        const double *gm = J.geom + NG * int64_t(valid ? e : 0);
#pragma unroll
        for (int u = 0; u < NG; ++u) L.gl[u] = gm[u];
#pragma unroll
        for (int j = 0; j < MFMA_SPCMAX; ++j)
        {
            const int q = 4 * j + kk;
            const bool in = valid && j < J.spc && q < J.nq;
            L.wq[j] = in ? J.qw[q] : 0.0;
#pragma unroll
            for (int u = 0; u < UMAX; ++u) L.cv[u][j] = in && u < T.nu ? T.cu[u][int64_t(e) * J.nq + q] : 0.0;
        }
    };
    for (int g = blockIdx.x * (MFMA_TPB / 64) + wave; g < ngroups; g += gridDim.x * (MFMA_TPB / 64))
    {
        // 1. this group's loads, all in flight together (loading the next group's during this
        //    one's matrix work was measured slower: the registers it holds cost a wave per SIMD)
        Loads cur;
        load(g, cur);
        const Jinv jc0 = jinv<NV>(cur.gl, 0.0, 0.0);
        // 2. per combo: its weights, then its spc steps of D (four points each) into the K loop,
        //    NT tiles of 16 local entries at a time
        for (int t0 = 0; t0 < J.ntile; t0 += NT)
        {
            f64x4 acc[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
            for (int c = 0; c < J.ncombo; ++c)
            {
                const int al = (J.alpha >> (2 * c)) & 3, be = (J.beta >> (2 * c)) & 3;
                double W[UMAX], W0 = 0.0;
                if constexpr (NV == 3) W0 = combo_weights<UMAX>(T, al, be, jc0, W);
                const double *Rc = R + (c * J.spc * 4 + kk) * npad + t0 * 16 + ci;
#pragma unroll
                for (int j = 0; j < MFMA_SPCMAX; ++j)
                {
                    if (j < J.spc)
                    {
                        double b[NT];
#pragma unroll
                        for (int t = 0; t < NT; ++t) b[t] = Rc[j * 4 * npad + t * 16];
                        double ad = jc0.ad;
                        if constexpr (NV == 4)
                        {
                            const int q = min(4 * j + kk, J.nq - 1);
                            const Jinv jq = jinv<NV>(cur.gl, J.qxi[q], J.qeta[q]);
                            W0 = combo_weights<UMAX>(T, al, be, jq, W);
                            ad = jq.ad;
                        }
                        double d = W0;
#pragma unroll
                        for (int u = 0; u < UMAX; ++u) d += cur.cv[u][j] * W[u];
                        d *= cur.wq[j] * ad;
#pragma unroll
                        for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(d, b[t], acc[t], 0, 0, 0);
                    }
                }
            }
            // 3. element v of this lane: cell 4 v + kk of the group, entry (t0 + t) * 16 + ci
#pragma unroll
            for (int t = 0; t < NT; ++t)
            {
                const int n = (t0 + t) * 16 + ci;
                const int2 oc = t0 == 0 ? oc1[t] : (n < J.nout ? J.ocol[n] : int2{-1, -1});
                if (oc.x < 0) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int cell = g * 16 + 4 * v + kk;
                    if (cell >= J.nc) continue;
                    double *Ke = J.E + cell * N2;
                    Ke[oc.x] = acc[t][v];
                    if (oc.y >= 0) Ke[oc.y] = acc[t][v];
                }
            }
        }
    }
}

// ---- 16-bit inputs, FP32 accumulation (the precision experiment) ----------------------------
// The same contraction on v_mfma_f32_16x16x16{f16,bf16}: D and R rounded to 16 bits, products
// summed in FP32, the element matrix widened to FP64.  One instruction takes 16 k: lane l holds
// A[cell l%16][k 4(l/16) .. 4(l/16)+3] and B[those k][entry l%16]; its result's element v is row
// (cell) 4 (l/16) + v, column (entry) l%16.  A combo's points are padded to steps of 16; R is laid
// out [step][entry][16 k] so that a lane reads its four k with one 8-byte load.
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even, to the upper 16 bits of the FP32 value
__device__ inline short to_bf16(float x)
{
    const unsigned u = __float_as_uint(x);
    return static_cast<short>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

constexpr int MFMA16_SPCMAX = 2;             // steps of 16 per combo: rules of at most 32 points

template <int NV, int NT, int UMAX, bool BF>
__global__ __launch_bounds__(MFMA_TPB) void k_elem_matrix_mfma16(MfmaJob J, MfmaTerms T)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ci = lane & 15, kk = lane >> 4;
    const int ngroups = (J.nc + 15) >> 4;
    const int64_t N2 = int64_t(J.N) * J.N;
    const int npad = 16 * J.ntile;
    const unsigned short *R16 = reinterpret_cast<const unsigned short *>(J.rt);
    constexpr int NG = NV == 4 ? 8 : 6;
    for (int g = blockIdx.x * (MFMA_TPB / 64) + wave; g < ngroups; g += gridDim.x * (MFMA_TPB / 64))
    {
        const int e = g * 16 + ci;
        const bool valid = e < J.nc;
        double gl[NG];
        const double *gm = J.geom + NG * int64_t(valid ? e : 0);
#pragma unroll
        for (int u = 0; u < NG; ++u) gl[u] = gm[u];
        // this lane's points: q = 16 j + 4 kk + r
        double wq[MFMA16_SPCMAX][4], cv[UMAX][MFMA16_SPCMAX][4];
#pragma unroll
        for (int j = 0; j < MFMA16_SPCMAX; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                const int q = 16 * j + 4 * kk + r;
                const bool in = valid && j < J.spc && q < J.nq;
                wq[j][r] = in ? J.qw[q] : 0.0;
#pragma unroll
                for (int u = 0; u < UMAX; ++u) cv[u][j][r] = in && u < T.nu ? T.cu[u][int64_t(e) * J.nq + q] : 0.0;
            }
        const Jinv jc0 = jinv<NV>(gl, 0.0, 0.0);
        for (int t0 = 0; t0 < J.ntile; t0 += NT)
        {
            f32x4 acc[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < J.ncombo; ++c)
            {
                const int al = (J.alpha >> (2 * c)) & 3, be = (J.beta >> (2 * c)) & 3;
                double W[UMAX], W0 = 0.0;
                if constexpr (NV == 3) W0 = combo_weights<UMAX>(T, al, be, jc0, W);
#pragma unroll
                for (int j = 0; j < MFMA16_SPCMAX; ++j)
                {
                    if (j >= J.spc) break;
                    const int step = c * J.spc + j;
                    f16x4 ah;
                    bf16x4 ab;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                    {
                        double ad = jc0.ad;
                        if constexpr (NV == 4)
                        {
                            const int q = min(16 * j + 4 * kk + r, J.nq - 1);
                            const Jinv jq = jinv<NV>(gl, J.qxi[q], J.qeta[q]);
                            W0 = combo_weights<UMAX>(T, al, be, jq, W);
                            ad = jq.ad;
                        }
                        double d = W0;
#pragma unroll
                        for (int u = 0; u < UMAX; ++u) d += cv[u][j][r] * W[u];
                        d *= wq[j][r] * ad;
                        if constexpr (BF) ab[r] = to_bf16(static_cast<float>(d));
                        else ah[r] = static_cast<_Float16>(static_cast<float>(d));
                    }
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                    {
                        const uint2 raw = *reinterpret_cast<const uint2 *>(
                            R16 + ((int64_t(step) * npad + (t0 + t) * 16 + ci) * 16 + 4 * kk));
                        if constexpr (BF)
                        {
                            bf16x4 b;
                            b[0] = short(raw.x & 0xffff); b[1] = short(raw.x >> 16);
                            b[2] = short(raw.y & 0xffff); b[3] = short(raw.y >> 16);
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ab, b, acc[t], 0, 0, 0);
                        }
                        else
                        {
                            const f16x4 b = __builtin_bit_cast(f16x4, raw);
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, b, acc[t], 0, 0, 0);
                        }
                    }
                }
            }
            // element v of this lane: cell 4 kk + v of the group, entry (t0 + t) * 16 + ci
#pragma unroll
            for (int t = 0; t < NT; ++t)
            {
                const int n = (t0 + t) * 16 + ci;
                if (n >= J.nout) continue;
                const int2 oc = J.ocol[n];
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int cell = g * 16 + 4 * kk + v;
                    if (cell >= J.nc) continue;
                    double *Ke = J.E + cell * N2;
                    Ke[oc.x] = static_cast<double>(acc[t][v]);
                    if (oc.y >= 0) Ke[oc.y] = static_cast<double>(acc[t][v]);
                }
            }
        }
    }
}

template <int NV, int NT, int UMAX>
void launch_u(const MfmaJob &J, const MfmaTerms &T, bool rlds, size_t lds, int grid, hipStream_t s)
{
    if (rlds)
        hipLaunchKernelGGL((k_elem_matrix_mfma<NV, NT, UMAX, true>), dim3(grid), dim3(MFMA_TPB), lds, s, J, T);
    else
        hipLaunchKernelGGL((k_elem_matrix_mfma<NV, NT, UMAX, false>), dim3(grid), dim3(MFMA_TPB), lds, s, J, T);
}

template <int NV, int NT>
void launch_nt(const MfmaJob &J, const MfmaTerms &T, bool rlds, size_t lds, int grid, hipStream_t s)
{
    if (T.nu <= 1) launch_u<NV, NT, 1>(J, T, rlds, lds, grid, s);
    else if (T.nu == 2) launch_u<NV, NT, 2>(J, T, rlds, lds, grid, s);
    else launch_u<NV, NT, MFMA_UMAX>(J, T, rlds, lds, grid, s);
}

template <int NV>
bool launch_nv(const MfmaJob &J, const MfmaTerms &T, int nt, bool rlds, size_t lds, int grid, hipStream_t s)
{
    switch (nt)
    {
    case 1: launch_nt<NV, 1>(J, T, rlds, lds, grid, s); return true;
    case 2: launch_nt<NV, 2>(J, T, rlds, lds, grid, s); return true;
    case 3: launch_nt<NV, 3>(J, T, rlds, lds, grid, s); return true;
    case 4: launch_nt<NV, 4>(J, T, rlds, lds, grid, s); return true;
    default: return false;
    }
}

// the terms by coefficient: constants folded into W0, the distinct arrays numbered; false past
// MFMA_UMAX arrays
bool mfma_terms(const TermSet &TS, MfmaTerms &T)
{
    T = MfmaTerms{};
    T.nt = TS.nt;
    for (int t = 0; t < TS.nt; ++t)
    {
        T.a[t] = TS.a[t];
        T.b[t] = TS.b[t];
        T.cval[t] = TS.cval[t];
        T.slot[t] = -1;
        if (!TS.c[t]) continue;
        for (int u = 0; u < T.nu && T.slot[t] < 0; ++u)
            if (T.cu[u] == TS.c[t]) T.slot[t] = u;
        if (T.slot[t] < 0)
        {
            if (T.nu == MFMA_UMAX) return false;
            T.cu[T.nu] = TS.c[t];
            T.slot[t] = T.nu++;
        }
    }
    return true;
}

template <int NV, int NT, bool BF>
void launch16_nt(const MfmaJob &J, const MfmaTerms &T, int grid, hipStream_t s)
{
    if (T.nu <= 1) hipLaunchKernelGGL((k_elem_matrix_mfma16<NV, NT, 1, BF>), dim3(grid), dim3(MFMA_TPB), 0, s, J, T);
    else if (T.nu == 2) hipLaunchKernelGGL((k_elem_matrix_mfma16<NV, NT, 2, BF>), dim3(grid), dim3(MFMA_TPB), 0, s, J, T);
    else hipLaunchKernelGGL((k_elem_matrix_mfma16<NV, NT, MFMA_UMAX, BF>), dim3(grid), dim3(MFMA_TPB), 0, s, J, T);
}

template <int NV, bool BF>
bool launch16_nv(const MfmaJob &J, const MfmaTerms &T, int nt, int grid, hipStream_t s)
{
    switch (nt)
    {
    case 1: launch16_nt<NV, 1, BF>(J, T, grid, s); return true;
    case 2: launch16_nt<NV, 2, BF>(J, T, grid, s); return true;
    case 4: launch16_nt<NV, 4, BF>(J, T, grid, s); return true;
    default: return false;
    }
}
} // namespace

size_t mfma_table_bytes(int ks, int ntile) { return sizeof(double) * size_t(4) * ks * 16 * ntile; }

bool launch_elem_matrix_mfma16(int nv, const MfmaJob &J, const TermSet &TS, int nt, bool bf16, hipStream_t s)
{
    if (J.nc == 0) return true;
    if (J.spc > MFMA16_SPCMAX || J.ntile % nt != 0) return false;
    MfmaTerms T;
    if (!mfma_terms(TS, T)) return false;
    const int grid = std::min(((J.nc + 15) / 16 + 3) / 4, 1024);
    if (nv == 4) return bf16 ? launch16_nv<4, true>(J, T, nt, grid, s) : launch16_nv<4, false>(J, T, nt, grid, s);
    return bf16 ? launch16_nv<3, true>(J, T, nt, grid, s) : launch16_nv<3, false>(J, T, nt, grid, s);
}

bool launch_elem_matrix_mfma(int nv, const MfmaJob &J, const TermSet &TS, int nt, bool rlds, hipStream_t s)
{
    if (J.nc == 0) return true;
    const size_t lds = rlds ? mfma_table_bytes(J.ks, J.ntile) : 0;
    if (lds > MFMA_LDS_MAX || J.spc > MFMA_SPCMAX || J.ntile % nt != 0) return false;
    MfmaTerms T;
    if (!mfma_terms(TS, T)) return false;
    const int groups = (J.nc + 15) / 16;
    const int grid = std::min((groups + 3) / 4, 1024);
    if (nv == 4) return launch_nv<4>(J, T, nt, rlds, lds, grid, s);
    return launch_nv<3>(J, T, nt, rlds, lds, grid, s);
}

} // namespace femd_amd
