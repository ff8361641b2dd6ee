// kernels.hip -- the gfx950 kernels of the 2D assemble -> SpMV -> Krylov path.
//
// Two families:
//  * exact kernels (FP64), which reproduce the reference's rounding sequence operation by
//    operation: this file is compiled with -ffp-contract=off, so a*b + c is two roundings, as
//    in the reference's x86-64 build; element entries are summed in ascending element order
//    (forms/assembly_2d.hpp:265-301), matrix rows left to right (sparse/csr.hpp:203-215), and
//    dot products in 4096-long chunks of eight interleaved lanes (util/omp.hpp:135-192);
//  * fast kernels (FP32 with FP64 accumulators, tree reductions) for the mixed-precision
//    inner solves and GMRES, whose results are checked against tolerances instead of bits.
//
// Wavefronts are 64 lanes; blocks are 256 threads (4 waves) unless noted.
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <cmath>

namespace femd_amd {

namespace {
constexpr int TPB = 256;

inline int grid_for(int64_t n, int tpb = TPB)
{
    int64_t g = (n + tpb - 1) / tpb;
    return static_cast<int>(g < 1 ? 1 : g);
}

__device__ inline double lane_combine(const double *s)
{
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// The chunk sum of util/omp.hpp:135-160 over prod[0, len) held in LDS: lane k adds the entries
// k, k+8, ... left to right; the entries after the last full group of eight go into lane 0;
// then the fixed combination.  Every thread of the block must call it; thread 0 gets the result.
__device__ inline double chunk_sum(const double *prod, int len, double *lane)
{
    const int tid = threadIdx.x, full = len & ~7;
    if (tid < 8)
    {
        // 16 reads in flight ahead of the 16 dependent additions, in order
        double s = 0.0;
        int i = tid;
        for (; i + 8 * 15 < full; i += 8 * 16)
        {
            double t[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) t[u] = prod[i + 8 * u];
#pragma unroll
            for (int u = 0; u < 16; ++u) s += t[u];
        }
        for (; i < full; i += 8) s += prod[i];
        lane[tid] = s;
    }
    __syncthreads();
    double out = 0.0;
    if (tid == 0)
    {
        double s0 = lane[0];
        for (int i = full; i < len; ++i) s0 += prod[i];
        double l[8] = {s0, lane[1], lane[2], lane[3], lane[4], lane[5], lane[6], lane[7]};
        out = lane_combine(l);
    }
    __syncthreads();
    return out;
}

// chunk_sum of two arrays at once: lanes 0-7 of the first wave take a's eight chains, lanes 8-15
// b's, in the same instructions; thread 0 gets a's sum in *sa, thread 8 b's in *sb (the same
// bits as two calls of chunk_sum).
__device__ inline void chunk_sum2(const double *a, const double *b, int len, double *lane, double *sa, double *sb)
{
    const int tid = threadIdx.x, full = len & ~7;
    if (tid < 16)
    {
        const double *src = tid < 8 ? a : b;
        double s = 0.0;
        int i = tid & 7;
        for (; i + 8 * 15 < full; i += 8 * 16)
        {
            double t[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) t[u] = src[i + 8 * u];
#pragma unroll
            for (int u = 0; u < 16; ++u) s += t[u];
        }
        for (; i < full; i += 8) s += src[i];
        lane[tid] = s;
    }
    __syncthreads();
    if (tid == 0 || tid == 8)
    {
        const double *src = tid == 0 ? a : b;
        const double *l0 = lane + tid;
        double s0 = l0[0];
        for (int i = full; i < len; ++i) s0 += src[i];
        double l[8] = {s0, l0[1], l0[2], l0[3], l0[4], l0[5], l0[6], l0[7]};
        *(tid == 0 ? sa : sb) = lane_combine(l);
    }
    __syncthreads();
}

// The chunk partial sums added left to right from zero (util/omp.hpp:186-191); one chunk is
// returned as it is.
// Computed by a whole workgroup (every thread must call it; thread 0 gets the result): the
// partials are staged through LDS in slices of TOTAL_SLICE by all threads, and thread 0 adds them
// left to right with many LDS reads in flight.
constexpr int TOTAL_SLICE = 2048;
__device__ inline double ordered_total_wg(const double *parts, int np, double *sh)
{
    if (np == 1) return parts[0];
    double s = 0.0;
    for (int c0 = 0; c0 < np; c0 += TOTAL_SLICE)
    {
        const int m = np - c0 < TOTAL_SLICE ? np - c0 : TOTAL_SLICE;
        for (int k = threadIdx.x; k < m; k += blockDim.x) sh[k] = parts[c0 + k];
        __syncthreads();
        if (threadIdx.x == 0)
        {
            int k = 0;
            for (; k + 15 < m; k += 16)
            {
                double t[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) t[u] = sh[k + u];
#pragma unroll
                for (int u = 0; u < 16; ++u) s += t[u];
            }
            for (; k < m; ++k) s += sh[k];
        }
        __syncthreads();
    }
    return s;
}
} // namespace

// ============================================================================ assembly
// Physical derivative row of code m at local function l (fe/cache_2d.hpp:211-226): the
// reference values, or J^{-T} times the reference gradient.
template <int N>
__device__ inline double phys(int m, const double *R, int l, double j0, double j1, double j2, double j3)
{
    if (m == 0) return R[l];
    if (m == 1) return j0 * R[N + l] + j2 * R[2 * N + l];
    return j1 * R[N + l] + j3 * R[2 * N + l];
}

// The Jacobian J = [dx/dxi, dx/deta, dy/dxi, dy/deta] of cell geometry g at (xi, eta): constant
// on a triangle (g = x0, y0, J), the bilinear map's on a quadrilateral (g = its four vertices,
// fe/lagrange_space_quad.hpp:163-169); then det, |det| and J^{-1} as the cache forms them
// (fe/cache_2d.hpp:106-115).
struct CellJ {
    double i0, i1, i2, i3, ad;
};
template <int NV>
__device__ inline CellJ cell_jacobian(const double *g, double xi, double eta)
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
    return CellJ{J3 / d, -J1 / d, -J2 / d, J0 / d, fabs(d)};
}

// One thread per (cell e, row i): the row Ke[i][:] of the element matrix, the terms in order, then
// the points, each product skipped when it is exactly zero (forms/assembly_2d.hpp:102-126).  NV
// is the cell's vertex count (3: P_k on triangles, J constant; 4: Q_k on quadrilaterals, J at
// every point).
template <int N, int NV>
__global__ __launch_bounds__(TPB) void k_elem_matrix(int nc, int nq, const double *__restrict__ geom,
                                                     const double *__restrict__ tab, const double *__restrict__ qw,
                                                     const double *__restrict__ qxi, const double *__restrict__ qeta,
                                                     TermSet T, double *__restrict__ E)
{
    extern __shared__ double sh[];
    double *stab = sh, *sqw = sh + nq * 3 * N;
    for (int k = threadIdx.x; k < nq * 3 * N; k += blockDim.x) stab[k] = tab[k];
    for (int k = threadIdx.x; k < nq; k += blockDim.x) sqw[k] = qw[k];
    __syncthreads();
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(nc) * N) return;
    const int e = static_cast<int>(idx / N), i = static_cast<int>(idx % N);
    const double *g = geom + (NV == 4 ? 8 : 6) * static_cast<int64_t>(e);
    CellJ cj = cell_jacobian<NV>(g, 0.0, 0.0);
    double K[N];
#pragma unroll
    for (int j = 0; j < N; ++j) K[j] = 0.0;
    for (int t = 0; t < T.nt; ++t)
    {
        const int ca = T.a[t], cb = T.b[t];
        const double *c = T.c[t];
        const double cv = T.cval[t];
        for (int q = 0; q < nq; ++q)
        {
            if constexpr (NV == 4) cj = cell_jacobian<NV>(g, qxi[q], qeta[q]);
            const double i0 = cj.i0, i1 = cj.i1, i2 = cj.i2, i3 = cj.i3;
            const double w = sqw[q] * cj.ad;
            const double s = w * (c ? c[static_cast<int64_t>(e) * nq + q] : cv);
            if (s == 0.0) continue;
            const double *R = stab + q * 3 * N;
            const double sa = s * phys<N>(ca, R, i, i0, i1, i2, i3);
            if (sa == 0.0) continue;
#pragma unroll
            for (int j = 0; j < N; ++j) K[j] += sa * phys<N>(cb, R, j, i0, i1, i2, i3);
        }
    }
    double *out = E + (static_cast<int64_t>(e) * N + i) * N;
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = K[j];
}

// One thread per (cell e, local l): fe[l] of assemble_vector_2d (forms/assembly_2d.hpp:184-196).
template <int N, int NV>
__global__ __launch_bounds__(TPB) void k_elem_vector(int nc, int nq, const double *__restrict__ geom,
                                                     const double *__restrict__ tab, const double *__restrict__ qw,
                                                     const double *__restrict__ qxi, const double *__restrict__ qeta,
                                                     TermSet T, double *__restrict__ F)
{
    extern __shared__ double sh[];
    double *stab = sh, *sqw = sh + nq * 3 * N;
    for (int k = threadIdx.x; k < nq * 3 * N; k += blockDim.x) stab[k] = tab[k];
    for (int k = threadIdx.x; k < nq; k += blockDim.x) sqw[k] = qw[k];
    __syncthreads();
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(nc) * N) return;
    const int e = static_cast<int>(idx / N), l = static_cast<int>(idx % N);
    const double *g = geom + (NV == 4 ? 8 : 6) * static_cast<int64_t>(e);
    CellJ cj = cell_jacobian<NV>(g, 0.0, 0.0);
    double f = 0.0;
    for (int t = 0; t < T.nt; ++t)
    {
        const double *c = T.c[t];
        const double cv = T.cval[t];
        for (int q = 0; q < nq; ++q)
        {
            if constexpr (NV == 4) cj = cell_jacobian<NV>(g, qxi[q], qeta[q]);
            const double i0 = cj.i0, i1 = cj.i1, i2 = cj.i2, i3 = cj.i3;
            const double w = sqw[q] * cj.ad;
            const double s = w * (c ? c[static_cast<int64_t>(e) * nq + q] : cv);
            if (s == 0.0) continue;
            f += s * phys<N>(T.a[t], stab + q * 3 * N, l, i0, i1, i2, i3);
        }
    }
    F[idx] = f;
}

// One thread per (boundary facet e, row i): the facet's element matrix as assemble_matrix_2d forms
// it on a facet cache (forms/assembly_2d.hpp:102-126 on fe/cache_2d.hpp:126-183): the cell's basis
// at the facet's points (the table of its local edge), J^{-1} and the weight of each point as the
// cache stores them; the loop order and zero skips of k_elem_matrix.
template <int N>
__global__ __launch_bounds__(TPB) void k_facet_matrix(int ne, int nq, const int *__restrict__ var, const double *__restrict__ w,
                                                      const double *__restrict__ Ji, const double *__restrict__ tab, TermSet T,
                                                      double *__restrict__ E)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(ne) * N) return;
    const int e = static_cast<int>(idx / N), i = static_cast<int>(idx % N);
    const double *tv = tab + static_cast<int64_t>(var[e]) * nq * 3 * N;
    double K[N];
#pragma unroll
    for (int j = 0; j < N; ++j) K[j] = 0.0;
    for (int t = 0; t < T.nt; ++t)
    {
        const int ca = T.a[t], cb = T.b[t];
        const double *c = T.c[t];
        const double cv = T.cval[t];
        for (int q = 0; q < nq; ++q)
        {
            const int64_t k = static_cast<int64_t>(e) * nq + q;
            const double i0 = Ji[4 * k], i1 = Ji[4 * k + 1], i2 = Ji[4 * k + 2], i3 = Ji[4 * k + 3];
            const double s = w[k] * (c ? c[k] : cv);
            if (s == 0.0) continue;
            const double *R = tv + q * 3 * N;
            const double sa = s * phys<N>(ca, R, i, i0, i1, i2, i3);
            if (sa == 0.0) continue;
#pragma unroll
            for (int j = 0; j < N; ++j) K[j] += sa * phys<N>(cb, R, j, i0, i1, i2, i3);
        }
    }
    double *out = E + (static_cast<int64_t>(e) * N + i) * N;
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = K[j];
}

// One thread per (boundary facet e, local l): the facet's element vector (assemble_vector_2d on a
// facet cache).
template <int N>
__global__ __launch_bounds__(TPB) void k_facet_vector(int ne, int nq, const int *__restrict__ var, const double *__restrict__ w,
                                                      const double *__restrict__ Ji, const double *__restrict__ tab, TermSet T,
                                                      double *__restrict__ F)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(ne) * N) return;
    const int e = static_cast<int>(idx / N), l = static_cast<int>(idx % N);
    const double *tv = tab + static_cast<int64_t>(var[e]) * nq * 3 * N;
    double f = 0.0;
    for (int t = 0; t < T.nt; ++t)
    {
        const double *c = T.c[t];
        const double cv = T.cval[t];
        for (int q = 0; q < nq; ++q)
        {
            const int64_t k = static_cast<int64_t>(e) * nq + q;
            const double i0 = Ji[4 * k], i1 = Ji[4 * k + 1], i2 = Ji[4 * k + 2], i3 = Ji[4 * k + 3];
            const double s = w[k] * (c ? c[k] : cv);
            if (s == 0.0) continue;
            f += s * phys<N>(T.a[t], tv + q * 3 * N, l, i0, i1, i2, i3);
        }
    }
    F[idx] = f;
}

// Slot k (dst index pos[k], or k without pos) receives the element-order sum of its
// contributions src[list[ptr[k]]], src[list[ptr[k]+1]], ...:
//   GATHER_SET      dst = ((0 + c0) + c1) + ...       (a matrix, or a group's own vector block)
//   GATHER_SET_PZ   dst = 0 + (((0 + c0) + c1) + ...) (the block added onto a zeroed vector)
//   GATHER_ADD      dst = dst + (((0 + c0) + c1) + ...) (a later group's vector block)
//   GATHER_CONTINUE dst = ((dst + c0) + c1) + ...     (a later group's matrix entries: every
//                   contribution is added onto the stored value, as assemble_matrix_2d does)
__global__ __launch_bounds__(TPB) void k_gather(int64_t n, const int *__restrict__ pos, const int *__restrict__ ptr,
                                                const int *__restrict__ list, const double *__restrict__ src,
                                                double *__restrict__ dst, int mode)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t p = pos ? pos[k] : k;
    double acc = mode == GATHER_CONTINUE ? dst[p] : 0.0;
    for (int q = ptr[k]; q < ptr[k + 1]; ++q) acc += src[list[q]];
    dst[p] = mode == GATHER_SET_PZ ? 0.0 + acc : mode == GATHER_ADD ? dst[p] + acc : acc;
}

// The Integrand register program (forms/integrand.hpp:300-342) at the quadrature points, the
// point mapped as the cache does: x = x0 + J00 xi + J01 eta on a triangle
// (fe/lagrange_space_2d.hpp:78-85), the bilinear map on a quadrilateral
// (fe/lagrange_space_quad.hpp:155-162).
// The library functions of the programs, called out of line: inlined into every case of the
// interpreter's switch they drive its register demand (and so its occupancy) for all programs.
__device__ __noinline__ double op_pow(double r, double n) { return pow(r, n); }
__device__ __noinline__ double op_sin(double r) { return sin(r); }
__device__ __noinline__ double op_cos(double r) { return cos(r); }
__device__ __noinline__ double op_exp(double r) { return exp(r); }
__device__ __noinline__ double op_log(double r) { return log(r); }
__device__ __noinline__ double op_tanh(double r) { return tanh(r); }
__device__ __noinline__ double op_sinh(double r) { return sinh(r); }
__device__ __noinline__ double op_cosh(double r) { return cosh(r); }

// The program's register file: an array of MAX_REGS per lane (scratch memory), or nreg rows of the
// workgroup's LDS, register i of lane t at [i * TPB + t] (consecutive lanes, consecutive banks).
template <bool LDS> struct RegFile;
template <> struct RegFile<false> {
    double r[MAX_REGS];
    __device__ double &operator[](int i) { return r[i]; }
};
template <> struct RegFile<true> {
    double *base;
    __device__ double &operator[](int i) { return base[i * TPB]; }
};

template <class RF>
__device__ inline double run_program(double x, double y, double nx, double ny, int ninstr, const int *__restrict__ code,
                                     const double *__restrict__ consts, RF &reg, const InputSet *in = nullptr, int64_t pt = 0)
{
    int res = 0;
    for (int s = 0; s < ninstr; ++s)
    {
        const int op = code[4 * s], o = code[4 * s + 1], a = code[4 * s + 2], b = code[4 * s + 3];
        double v;
        switch (op)
        {
        case 0: v = consts[a]; break;
        case 1: v = x; break;
        case 2: v = y; break;
        case 3: v = (a < 0 ? -1.0 : 1.0) * nx; break;
        case 4: v = (a < 0 ? -1.0 : 1.0) * ny; break;
        case 5: v = in->p[a][pt]; break;
        case 6: v = reg[a] + reg[b]; break;
        case 7: v = reg[a] * reg[b]; break;
        case 8:
        {
            const double n = consts[b], r = reg[a];
            v = n == 2.0 ? r * r : n == -1.0 ? 1.0 / r : n == 0.5 ? sqrt(r) : op_pow(r, n);
            break;
        }
        case 9: v = op_sin(reg[a]); break;
        case 10: v = op_cos(reg[a]); break;
        case 11: v = op_exp(reg[a]); break;
        case 12: v = op_log(reg[a]); break;
        case 13: v = op_tanh(reg[a]); break;
        case 14: v = sqrt(reg[a]); break;
        case 15: v = op_sinh(reg[a]); break;
        case 16: v = op_cosh(reg[a]); break;
        case 17: { const double t = op_exp(-fabs(reg[a])); v = 2.0 * t / (1.0 + t * t); break; }
        case 18: v = fabs(reg[a]); break;
        case 19: v = reg[a] > 0.0 ? 1.0 : (reg[a] < 0.0 ? -1.0 : 0.0); break;
        default: v = 0.0; break;
        }
        reg[o] = v;
        res = o;
    }
    return reg[res];
}

// The program at given points (boundary facets: x, y and the outward normal per point).
__global__ __launch_bounds__(TPB) void k_coefficient_pts(int64_t np, const double *__restrict__ xs,
                                                         const double *__restrict__ ys, const double *__restrict__ nxs,
                                                         const double *__restrict__ nys, int ninstr,
                                                         const int *__restrict__ code, const double *__restrict__ consts,
                                                         double *__restrict__ out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= np) return;
    RegFile<false> rf;
    out[k] = run_program(xs[k], ys[k], nxs[k], nys[k], ninstr, code, consts, rf);
}

// LDS: the register file in dynamic LDS (nreg * TPB doubles), else in scratch memory
template <bool LDS>
__global__ __launch_bounds__(TPB) void k_coefficient(int nc, int nq, int nv, const double *__restrict__ geom,
                                                     const double *__restrict__ qxi, const double *__restrict__ qeta,
                                                     int ninstr, const int *__restrict__ code, int nreg,
                                                     const double *__restrict__ consts, InputSet in, double *__restrict__ out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= static_cast<int64_t>(nc) * nq) return;
    const int e = static_cast<int>(k / nq), q = static_cast<int>(k % nq);
    double x, y;
    if (nv == 4)
    {
        const double *g = geom + 8 * static_cast<int64_t>(e);
        const double xi = qxi[q], eta = qeta[q];
        const double a0 = (1 - xi) * (1 - eta), a1 = xi * (1 - eta), a2 = xi * eta, a3 = (1 - xi) * eta;
        x = a0 * g[0] + a1 * g[2] + a2 * g[4] + a3 * g[6];
        y = a0 * g[1] + a1 * g[3] + a2 * g[5] + a3 * g[7];
    }
    else
    {
        const double *g = geom + 6 * static_cast<int64_t>(e);
        x = g[0] + g[2] * qxi[q] + g[3] * qeta[q];
        y = g[1] + g[4] * qxi[q] + g[5] * qeta[q];
    }
    (void)nreg;
    if constexpr (LDS)
    {
        extern __shared__ double sh[];
        RegFile<true> rf{sh + threadIdx.x};
        out[k] = run_program(x, y, 0.0, 0.0, ninstr, code, consts, rf, &in, k);
    }
    else
    {
        RegFile<false> rf;
        out[k] = run_program(x, y, 0.0, 0.0, ninstr, code, consts, rf, &in, k);
    }
}

// One thread per (cell e, point q): a known field's value or x / y derivative there, the sums of
// Cache2D::eval_multi (fe/cache_2d.hpp:245-288) in its order: the cell's coefficients gathered
// (eliminated nodes read 0), then over the local functions l, value R[l] loc[l], or the row
// a R_xi[l] + b R_eta[l] (J^{-1}'s column for x or y, as the cache forms it) times loc[l].
template <int N, int NV>
__global__ __launch_bounds__(TPB) void k_eval_field(int nc, int nq, const int *__restrict__ cell_dof,
                                                    const double *__restrict__ u, const double *__restrict__ tab,
                                                    const double *__restrict__ geom, const double *__restrict__ qxi,
                                                    const double *__restrict__ qeta, int code, double *__restrict__ out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= static_cast<int64_t>(nc) * nq) return;
    const int e = static_cast<int>(k / nq), q = static_cast<int>(k % nq);
    const int *d = cell_dof + static_cast<int64_t>(e) * N;
    double loc[N];
#pragma unroll
    for (int l = 0; l < N; ++l) loc[l] = d[l] >= 0 ? u[d[l]] : 0.0;
    const double *R = tab + static_cast<int64_t>(q) * 3 * N;
    double s = 0.0;
    if (code == 0)
    {
#pragma unroll
        for (int l = 0; l < N; ++l) s += R[l] * loc[l];
    }
    else
    {
        const CellJ cj = cell_jacobian<NV>(geom + (NV == 4 ? 8 : 6) * static_cast<int64_t>(e), qxi[q], qeta[q]);
        const double a = code == 1 ? cj.i0 : cj.i1, b = code == 1 ? cj.i2 : cj.i3;
#pragma unroll
        for (int l = 0; l < N; ++l)
        {
            const double r = a * R[N + l] + b * R[2 * N + l];
            s += r * loc[l];
        }
    }
    out[k] = s;
}

// ============================================================================ SpMV
// ============================================================================ ordered dot
__global__ __launch_bounds__(TPB) void k_dot_parts(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                                                   double *__restrict__ parts, const int *stop)
{
    if (stop && *stop) return;
    __shared__ double prod[CHUNK];
    __shared__ double lane[8];
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * CHUNK;
    const int len = static_cast<int>(n - lo < CHUNK ? n - lo : CHUNK);
    // every load issued at once, on an index inside the chunk (a load under a condition, or one
    // per trip of a loop, waits for the loads before it)
    double a[CHUNK / TPB], b[CHUNK / TPB];
#pragma unroll
    for (int j = 0; j < CHUNK / TPB; ++j)
    {
        const int k = j * TPB + static_cast<int>(threadIdx.x), kc = k < len ? k : 0;
        a[j] = x[lo + kc];
        b[j] = y[lo + kc];
    }
#pragma unroll
    for (int j = 0; j < CHUNK / TPB; ++j)
    {
        const int k = j * TPB + static_cast<int>(threadIdx.x);
        if (k < len) prod[k] = a[j] * b[j];
    }
    __syncthreads();
    const double s = chunk_sum(prod, len, lane);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ __launch_bounds__(TPB) void k_sum_parts(int np, const double *parts, double *out)
{
    __shared__ double sh[TOTAL_SLICE];
    const double t = np == 0 ? 0.0 : ordered_total_wg(parts, np, sh);
    if (threadIdx.x == 0) out[0] = t;
}

// ============================================================================ vectors
__global__ void k_sub(int64_t n, const double *b, const double *a, double *r)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - a[i];
}
__global__ void k_mul(int64_t n, const double *a, const double *b, double *c)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] * b[i];
}
__global__ void k_copy(int64_t n, const double *a, double *b)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}
__global__ void k_fill(int64_t n, double v, double *a)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_gather_perm(int n, const int *perm, const double *src, double *dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}
__global__ void k_scatter_perm(int n, const int *perm, const double *src, double *dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[perm[k]] = src[k];
}
__global__ void k_take(int64_t n, const int *idx, const double *src, double *dst)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_to_float(int64_t n, const double *a, float *b)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = static_cast<float>(a[i]);
}
// Position of the diagonal in each sorted row, -1 when absent.
__global__ void k_diag_positions(int n, const int *ip, const int *ix, int *diag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = ip[i], hi = ip[i + 1];
    while (lo < hi) { const int m = (lo + hi) >> 1; if (ix[m] < i) lo = m + 1; else hi = m; }
    diag[i] = (lo < ip[i + 1] && ix[lo] == i) ? lo : -1;
}
// inv_i = 1 / a_ii (precond.hpp:88-92); the first row with no or a zero diagonal is reported.
__global__ void k_inv_diag(int n, const int *diag, const double *v, double *inv, int *bad_row)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = diag[i];
    if (d < 0 || v[d] == 0.0) { atomicMin(bad_row, i); if (inv) inv[i] = 0.0; return; }
    if (inv) inv[i] = 1.0 / v[d];
}
// max |a_ij| and max |a_ij - a_ji| (sparse/csr.hpp:631-644); maxima are order independent.
// The position of every entry's transpose (j, i) in row j, or -1 (a binary search in the sorted
// row; once per pattern).
// sixteen consecutive lanes per row i: entry p = (i, j) looks for (j, i) by a binary search of row j
constexpr int TP_LANES = 16;
__global__ __launch_bounds__(TPB) void k_transpose_positions(int n, const int *ip, const int *ix, int *tpos)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = static_cast<int>(t / TP_LANES);
    if (i >= n) return;
    for (int p = ip[i] + static_cast<int>(t % TP_LANES); p < ip[i + 1]; p += TP_LANES)
    {
        const int j = ix[p];
        int lo = ip[j], hi = ip[j + 1];
        while (lo < hi) { const int m = (lo + hi) >> 1; if (ix[m] < i) lo = m + 1; else hi = m; }
        tpos[p] = (lo < ip[j + 1] && ix[lo] == i) ? lo : -1;
    }
}

// max |a_ij| and max |a_ij - a_ji| (a missing transpose counts as 0), one pass over the entries
// with the pattern's transpose positions
__global__ __launch_bounds__(TPB) void k_asymmetry(int64_t nnz, const int *__restrict__ tpos, const double *__restrict__ v,
                                                   double *out2)
{
    __shared__ double sb[TPB], sd[TPB];
    double big = 0.0, dif = 0.0;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x)
    {
        const double a = v[p];
        const int q = tpos[p];
        big = fmax(big, fabs(a));
        // |a_ij - a_ji| once per pair, from the entry that comes first (the gather of the other
        // half is skipped); an entry without a transpose against 0
        if (q < 0) dif = fmax(dif, fabs(a));
        else if (q > p) dif = fmax(dif, fabs(a - v[q]));
    }
    sb[threadIdx.x] = big; sd[threadIdx.x] = dif;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) { sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + s]); sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + s]); }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        // non-negative doubles order like their bit patterns: an integer max is exact
        atomicMax(reinterpret_cast<unsigned long long *>(out2), __double_as_longlong(sb[0]));
        atomicMax(reinterpret_cast<unsigned long long *>(out2 + 1), __double_as_longlong(sd[0]));
    }
}

// ============================================================================ CG (krylov/cg.hpp)
// alpha = rz / p^T A p, with the breakdown test of cg.hpp:271-273 / 328-329.
__global__ __launch_bounds__(TPB) void k_cg_alpha(const double *parts, int np, CGState *st)
{
    if (st->stop) return;
    __shared__ double sh[TOTAL_SLICE];
    const double pAp = ordered_total_wg(parts, np, sh);
    if (threadIdx.x != 0) return;
    st->pAp = pAp;
    if (!(pAp > 0.0)) { st->breakdown = 1; st->stop = 1; return; }
    st->alpha = st->rz / pAp;
}

// x += alpha p, r -= alpha Ap over one chunk, and the chunk's sums of r_i^2 and r_i^2 inv_i
// (cg.hpp:33-102, 105-171).
__global__ __launch_bounds__(TPB) void k_cg_update(int64_t n, double *__restrict__ x, double *__restrict__ r,
                                                   const double *__restrict__ p, const double *__restrict__ Ap,
                                                   const double *__restrict__ inv, double *__restrict__ parts_rr,
                                                   double *__restrict__ parts_rz, const CGState *st)
{
    if (st->stop) return;
    __shared__ double prod[CHUNK];
    __shared__ double prod2[CHUNK];
    __shared__ double lane[16];
    const double alpha = st->alpha;
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * CHUNK;
    const int len = static_cast<int>(n - lo < CHUNK ? n - lo : CHUNK);
    // every load issued at once, on an index inside the chunk (a load under a condition waits for
    // the loads before it); the stores keep to the chunk
    double rv[CHUNK / TPB], av[CHUNK / TPB], xv[CHUNK / TPB], pv[CHUNK / TPB], dv[CHUNK / TPB];
#pragma unroll
    for (int j = 0; j < CHUNK / TPB; ++j)
    {
        const int k = j * TPB + static_cast<int>(threadIdx.x), kc = k < len ? k : 0;
        rv[j] = r[lo + kc];
        av[j] = Ap[lo + kc];
        xv[j] = x[lo + kc];
        pv[j] = p[lo + kc];
        dv[j] = inv ? inv[lo + kc] : 0.0;
    }
    double e2[CHUNK / TPB];
#pragma unroll
    for (int j = 0; j < CHUNK / TPB; ++j)
    {
        const int k = j * TPB + static_cast<int>(threadIdx.x);
        const double ri = rv[j] - alpha * av[j];
        e2[j] = ri * ri;
        if (k < len)
        {
            const int64_t i = lo + k;
            r[i] = ri;
            x[i] = xv[j] + alpha * pv[j];
            prod[k] = e2[j];
        }
    }
    if (!inv)
    {
        __syncthreads();
        const double srr = chunk_sum(prod, len, lane);
        if (threadIdx.x == 0) parts_rr[blockIdx.x] = srr;
        return;
    }
#pragma unroll
    for (int j = 0; j < CHUNK / TPB; ++j)
    {
        const int k = j * TPB + static_cast<int>(threadIdx.x);
        if (k < len) prod2[k] = e2[j] * dv[j];
    }
    __syncthreads();
    // r.r and r.D^-1 r of the chunk, their chains side by side
    double srr = 0.0, srz = 0.0;
    chunk_sum2(prod, prod2, len, lane, &srr, &srz);
    if (threadIdx.x == 0) parts_rr[blockIdx.x] = srr;
    if (threadIdx.x == 8) parts_rz[blockIdx.x] = srz;
}

// After the update: ||r||, the stopping test, beta and the loop-top tests (cg.hpp:330-349).
// jacobi_fused: r^T D^{-1} r came with the update (cg_jacobi); else rz is the next step's job
// (general preconditioner: k_cg_rz after z = M^{-1} r), or rz = rr without a preconditioner.
__global__ __launch_bounds__(TPB) void k_cg_beta(const double *parts_rr, const double *parts_rz, int np, int mode, CGState *st)
{
    if (st->stop) return;
    __shared__ double sh[TOTAL_SLICE];
    const double rr = ordered_total_wg(parts_rr, np, sh);
    const double rz1 = mode == 1 ? ordered_total_wg(parts_rz, np, sh) : rr;
    if (threadIdx.x != 0) return;
    st->rr = rr;
    st->rn = sqrt(rr);
    st->it += 1;
    if (st->rn <= st->tolnb) { st->stop = 1; st->converged_break = 1; return; }
    if (mode == 2) return;                                   // general M: beta after z = M^{-1} r
    const double rzn = rz1;
    st->beta = rzn / st->rz;
    st->rz = rzn;
    if (st->it >= st->max_iter) st->stop = 1;
    else if (!(st->rz > 0.0)) { st->breakdown = 1; st->stop = 1; }
}

// General preconditioner: rz_new = r^T z, beta, and the loop-top tests (cg.hpp:277-283).
__global__ __launch_bounds__(TPB) void k_cg_rz(const double *parts, int np, CGState *st)
{
    if (st->stop) return;
    __shared__ double sh[TOTAL_SLICE];
    const double rzn = ordered_total_wg(parts, np, sh);
    if (threadIdx.x != 0) return;
    st->beta = rzn / st->rz;
    st->rz = rzn;
    if (st->it >= st->max_iter) st->stop = 1;
    else if (!(st->rz > 0.0)) { st->breakdown = 1; st->stop = 1; }
}

// p = inv r + beta p (cg.hpp:340-341), or p = r + beta p when inv is null (also p = z + beta p).
__global__ void k_cg_pupdate(int64_t n, double *__restrict__ p, const double *__restrict__ r,
                             const double *__restrict__ inv, const CGState *st)
{
    if (st->stop) return;
    const double beta = st->beta;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = inv ? inv[i] * r[i] + beta * p[i] : r[i] + beta * p[i];
}

// ============================================================================ ILU(0) / SSOR
// One workgroup walks the level sets of the sweep; rows of a level run in parallel, each row with
// the reference's own operation order, and a barrier separates the levels.  The whole sweep stays
// inside one workgroup, so no memory traffic needs cross-CU ordering.
constexpr int WG_SWEEP = 1024;

// IKJ ILU(0) on the pattern (precond.hpp:158-182), row by row within each level.
__global__ __launch_bounds__(WG_SWEEP) void k_ilu0_factor(int nlev, const int *__restrict__ lptr,
                                                          const int *__restrict__ lrows, const int *__restrict__ ip,
                                                          const int *__restrict__ ix, const int *__restrict__ diag,
                                                          double *v, int *bad_row)
{
    for (int l = 0; l < nlev; ++l)
    {
        for (int k = lptr[l] + threadIdx.x; k < lptr[l + 1]; k += blockDim.x)
        {
            const int i = lrows[k], re = ip[i + 1], di = diag[i];
            for (int p = ip[i]; p < di; ++p)
            {
                const int kk = ix[p];
                const double lik = v[p] / v[diag[kk]];
                v[p] = lik;
                int t = p + 1;
                for (int q = diag[kk] + 1; q < ip[kk + 1]; ++q)
                {
                    const int col = ix[q];
                    while (t < re && ix[t] < col) ++t;
                    if (t == re) break;
                    if (ix[t] == col) v[t] = v[t] - lik * v[q];
                }
            }
            if (v[di] == 0.0) atomicMin(bad_row, i);
        }
        __syncthreads();
    }
}

// IKJ ILU(0) (precond.hpp:158-182) for the rows of one dependency level, one wave per row, every
// wave of the GPU at work: the row's entries, the pivots and upper parts of the rows it eliminates
// with, and where each of their columns sits in the row, are staged in LDS first (all loads in
// flight together); then the eliminations run in the reference's order -- for each lower entry k
// in stored order, l = a_ik / u_kk, then a_ij = a_ij - l u_kj for every j of row k's upper part in
// row i -- so every value is the reference's bit for bit.  Rows of at most ILU_ROW entries whose
// eliminations touch at most ILU_UPD upper entries in all (the host checks; else k_ilu0_factor).
constexpr int ILU_ROW = 64, ILU_UPD = 1024, ILU_WAVES = 4;

__global__ __launch_bounds__(64 * ILU_WAVES) void k_ilu0_level(int nrows, const int *__restrict__ rows,
                                                               const int *__restrict__ ip, const int *__restrict__ ix,
                                                               const int *__restrict__ diag, double *v, int *bad_row)
{
    struct Stage {
        int cols[ILU_ROW];
        int uoff[ILU_ROW + 1];
        int udiag[ILU_ROW];
        double vals[ILU_ROW];
        double ukk[ILU_ROW];
        int ucol[ILU_UPD];
        double uval[ILU_UPD];
        signed char at[ILU_UPD];
    };
    __shared__ Stage st_all[ILU_WAVES];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * ILU_WAVES + wv;
    if (r >= nrows) return;
    Stage &S = st_all[wv];
    const int i = rows[r], p0 = ip[i], len = ip[i + 1] - p0, nlow = diag[i] - p0;
    auto wave_sync = [] { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); __builtin_amdgcn_wave_barrier(); };
    // the row, and per lower entry k its pivot and where its upper part starts and ends
    int cnt = 0;
    if (lane < len)
    {
        S.cols[lane] = ix[p0 + lane];
        S.vals[lane] = v[p0 + lane];
    }
    if (lane < nlow)
    {
        const int k = ix[p0 + lane], dk = diag[k];
        S.udiag[lane] = dk;
        S.ukk[lane] = v[dk];
        cnt = ip[k + 1] - dk - 1;
    }
    // exclusive prefix of the upper parts' sizes over the lower entries
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane < nlow) S.uoff[lane] = incl - cnt;
    if (lane == nlow - 1) S.uoff[nlow] = incl;
    wave_sync();
    const int total = nlow > 0 ? S.uoff[nlow] : 0;
    // the upper parts, and each entry's place in row i past the lower entry it comes from (-1: none)
    for (int q = lane; q < total; q += 64)
    {
        int lo = 0, hi = nlow - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (S.uoff[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        const int at = S.udiag[lo] + 1 + (q - S.uoff[lo]);
        S.ucol[q] = ix[at];
        S.uval[q] = v[at];
    }
    wave_sync();
    for (int q = lane; q < total; q += 64)
    {
        int lo = 0, hi = nlow - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (S.uoff[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        const int c = S.ucol[q];
        int a = lo + 1, b = len;
        while (a < b)
        {
            const int mid = (a + b) >> 1;
            if (S.cols[mid] < c) a = mid + 1;
            else b = mid;
        }
        S.at[q] = static_cast<signed char>(a < len && S.cols[a] == c ? a : -1);
    }
    wave_sync();
    // the eliminations, lower entry by lower entry
    for (int t = 0; t < nlow; ++t)
    {
        const double l = S.vals[t] / S.ukk[t];
        wave_sync();
        if (lane == 0) S.vals[t] = l;
        const int q = S.uoff[t] + lane;
        if (q < S.uoff[t + 1])
        {
            const int j = S.at[q];
            if (j >= 0) S.vals[j] = S.vals[j] - l * S.uval[q];
        }
        wave_sync();
    }
    if (lane < len) v[p0 + lane] = S.vals[lane];
    if (lane == 0 && S.vals[nlow] == 0.0) atomicMin(bad_row, i);
}

// the smallest pivot u_ii (a minimum: any order of the comparisons gives the same value)
__global__ __launch_bounds__(1024) void k_min_pivot(int n, const int *__restrict__ diag, const double *__restrict__ v,
                                                    double *out)
{
    __shared__ double part[1024];
    double m = __builtin_inf();
    for (int i = threadIdx.x; i < n; i += 1024) m = fmin(m, v[diag[i]]);
    part[threadIdx.x] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1)
    {
        if (int(threadIdx.x) < w) part[threadIdx.x] = fmin(part[threadIdx.x], part[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// z = U^{-1} L^{-1} r (precond.hpp:189-205).
__global__ __launch_bounds__(WG_SWEEP) void k_ilu0_apply(int nlevL, const int *__restrict__ lptrL,
                                                         const int *__restrict__ lrowsL, int nlevU,
                                                         const int *__restrict__ lptrU, const int *__restrict__ lrowsU,
                                                         const int *__restrict__ ip, const int *__restrict__ ix,
                                                         const int *__restrict__ diag, const double *__restrict__ v,
                                                         const double *__restrict__ r, double *z, const int *stop)
{
    if (stop && *stop) return;
    for (int l = 0; l < nlevL; ++l)
    {
        for (int k = lptrL[l] + threadIdx.x; k < lptrL[l + 1]; k += blockDim.x)
        {
            const int i = lrowsL[k];
            double s = r[i];
            for (int p = ip[i]; p < diag[i]; ++p) s = s - v[p] * z[ix[p]];
            z[i] = s;
        }
        __syncthreads();
    }
    for (int l = 0; l < nlevU; ++l)
    {
        for (int k = lptrU[l] + threadIdx.x; k < lptrU[l + 1]; k += blockDim.x)
        {
            const int i = lrowsU[k], di = diag[i];
            double s = z[i];
            for (int p = di + 1; p < ip[i + 1]; ++p) s = s - v[p] * z[ix[p]];
            z[i] = s / v[di];
        }
        __syncthreads();
    }
}

// SSOR(omega) (precond.hpp:125-144): forward sweep, diagonal scaling, backward sweep.
__global__ __launch_bounds__(WG_SWEEP) void k_ssor_apply(int n, int nlevL, const int *__restrict__ lptrL,
                                                         const int *__restrict__ lrowsL, int nlevU,
                                                         const int *__restrict__ lptrU, const int *__restrict__ lrowsU,
                                                         const int *__restrict__ ip, const int *__restrict__ ix,
                                                         const int *__restrict__ diag, const double *__restrict__ v,
                                                         double w, const double *__restrict__ r, double *z, const int *stop)
{
    if (stop && *stop) return;
    for (int l = 0; l < nlevL; ++l)
    {
        for (int k = lptrL[l] + threadIdx.x; k < lptrL[l + 1]; k += blockDim.x)
        {
            const int i = lrowsL[k];
            double s = r[i];
            for (int p = ip[i]; p < diag[i]; ++p) s = s - v[p] * z[ix[p]];
            z[i] = s * w / v[diag[i]];
        }
        __syncthreads();
    }
    const double c = (2.0 - w) / w;
    for (int i = threadIdx.x; i < n; i += blockDim.x) z[i] = z[i] * (c * v[diag[i]] / w);
    __syncthreads();
    for (int l = 0; l < nlevU; ++l)
    {
        for (int k = lptrU[l] + threadIdx.x; k < lptrU[l + 1]; k += blockDim.x)
        {
            const int i = lrowsU[k], di = diag[i];
            double s = z[i];
            for (int p = di + 1; p < ip[i + 1]; ++p) s = s - v[p] * z[ix[p]];
            z[i] = s * w / v[di];
        }
        __syncthreads();
    }
}

// ============================================================================ fast reductions
constexpr int FAST_MAX_BLOCKS = 1024;

__device__ inline double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double out = sh[0];
    __syncthreads();
    return out;
}

__global__ __launch_bounds__(TPB) void k_fdot(int64_t n, const double *x, const double *y, double *parts)
{
    __shared__ double sh[TPB];
    double s = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * TPB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * TPB)
        s = fma(x[i], y[i], s);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ __launch_bounds__(TPB) void k_fsum(int np, const double *parts, double *out)
{
    __shared__ double sh[TPB];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += TPB) s += parts[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[0] = s;
}

__global__ void k_axpy_dev(int64_t n, const double *alpha, double sign, const double *x, double *y)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += (sign * alpha[0]) * x[i];
}
__global__ void k_scale_dev(int64_t n, const double *num, const double *den, double *x)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= (num ? num[0] : 1.0) / den[0];
}
__global__ void k_axpy(int64_t n, double a, const double *x, double *y)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
__global__ void k_scale(int64_t n, double a, double *x)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= a;
}
struct VecPtrs { const double *p[64]; };
// x += sum_j y[j] V[j] (the solution update of csnewton.hpp:147-156), y on the device.
__global__ void k_lincomb(int64_t n, int k, const double *y, VecPtrs V, double *x)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = x[i];
    for (int j = 0; j < k; ++j) s += y[j] * V.p[j][i];
    x[i] = s;
}

// ============================================================================ launchers
#define FEMD_LAUNCH(kernel, grid, block, shmem, stream, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream, __VA_ARGS__)

#define ELEM_DISPATCH(KERNEL, N_, NV_)                                                              \
    case N_:                                                                                         \
        FEMD_LAUNCH((KERNEL<N_, NV_>), grid_for(static_cast<int64_t>(nc) * N_), TPB, sh, s, nc, nq, geom, tab, qw, qxi, \
                    qeta, T, out);                                                                   \
        break;

// P1..P5 on triangles (nloc 3, 6, 10, 15, 21) and Q1..Q4 on quadrilaterals (4, 9, 16, 25): the
// local sizes do not collide, and nv says which cell it is.
bool launch_elem_matrix(int nv, int nloc, int nc, int nq, const double *geom, const double *tab, const double *qw,
                        const double *qxi, const double *qeta, const TermSet &T, double *out, hipStream_t s)
{
    if (nc == 0) return true;
    const size_t sh = sizeof(double) * (static_cast<size_t>(nq) * 3 * nloc + nq);
    if (sh > 65536) return false;
    if (nv == 4)
        switch (nloc)
        {
            ELEM_DISPATCH(k_elem_matrix, 4, 4)
            ELEM_DISPATCH(k_elem_matrix, 9, 4)
            ELEM_DISPATCH(k_elem_matrix, 16, 4)
            ELEM_DISPATCH(k_elem_matrix, 25, 4)
        default: return false;
        }
    else
        switch (nloc)
        {
            ELEM_DISPATCH(k_elem_matrix, 3, 3)
            ELEM_DISPATCH(k_elem_matrix, 6, 3)
            ELEM_DISPATCH(k_elem_matrix, 10, 3)
            ELEM_DISPATCH(k_elem_matrix, 15, 3)
            ELEM_DISPATCH(k_elem_matrix, 21, 3)
        default: return false;
        }
    return true;
}

bool launch_elem_vector(int nv, int nloc, int nc, int nq, const double *geom, const double *tab, const double *qw,
                        const double *qxi, const double *qeta, const TermSet &T, double *out, hipStream_t s)
{
    if (nc == 0) return true;
    const size_t sh = sizeof(double) * (static_cast<size_t>(nq) * 3 * nloc + nq);
    if (sh > 65536) return false;
    if (nv == 4)
        switch (nloc)
        {
            ELEM_DISPATCH(k_elem_vector, 4, 4)
            ELEM_DISPATCH(k_elem_vector, 9, 4)
            ELEM_DISPATCH(k_elem_vector, 16, 4)
            ELEM_DISPATCH(k_elem_vector, 25, 4)
        default: return false;
        }
    else
        switch (nloc)
        {
            ELEM_DISPATCH(k_elem_vector, 3, 3)
            ELEM_DISPATCH(k_elem_vector, 6, 3)
            ELEM_DISPATCH(k_elem_vector, 10, 3)
            ELEM_DISPATCH(k_elem_vector, 15, 3)
            ELEM_DISPATCH(k_elem_vector, 21, 3)
        default: return false;
        }
    return true;
}

void launch_gather(int64_t n, const int *ptr, const int *list, const double *src, double *dst, bool plus_zero, hipStream_t s)
{
    if (n > 0)
        FEMD_LAUNCH(k_gather, grid_for(n), TPB, 0, s, n, static_cast<const int *>(nullptr), ptr, list, src, dst,
                    plus_zero ? GATHER_SET_PZ : GATHER_SET);
}

void launch_gather_mode(int64_t n, const int *pos, const int *ptr, const int *list, const double *src, double *dst, int mode,
                        hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_gather, grid_for(n), TPB, 0, s, n, pos, ptr, list, src, dst, mode);
}

#define FACET_DISPATCH(KERNEL, N_)                                                                    \
    case N_:                                                                                         \
        FEMD_LAUNCH((KERNEL<N_>), grid_for(static_cast<int64_t>(ne) * N_), TPB, 0, s, ne, nq, var, w, Ji, tab, T, out); \
        break;

bool launch_facet_matrix(int nloc, int ne, int nq, const int *var, const double *w, const double *Ji, const double *tab,
                         const TermSet &T, double *out, hipStream_t s)
{
    if (ne == 0) return true;
    switch (nloc)
    {
        FACET_DISPATCH(k_facet_matrix, 3)
        FACET_DISPATCH(k_facet_matrix, 4)
        FACET_DISPATCH(k_facet_matrix, 6)
        FACET_DISPATCH(k_facet_matrix, 9)
        FACET_DISPATCH(k_facet_matrix, 10)
        FACET_DISPATCH(k_facet_matrix, 15)
        FACET_DISPATCH(k_facet_matrix, 16)
        FACET_DISPATCH(k_facet_matrix, 21)
        FACET_DISPATCH(k_facet_matrix, 25)
    default: return false;
    }
    return true;
}

bool launch_facet_vector(int nloc, int ne, int nq, const int *var, const double *w, const double *Ji, const double *tab,
                         const TermSet &T, double *out, hipStream_t s)
{
    if (ne == 0) return true;
    switch (nloc)
    {
        FACET_DISPATCH(k_facet_vector, 3)
        FACET_DISPATCH(k_facet_vector, 4)
        FACET_DISPATCH(k_facet_vector, 6)
        FACET_DISPATCH(k_facet_vector, 9)
        FACET_DISPATCH(k_facet_vector, 10)
        FACET_DISPATCH(k_facet_vector, 15)
        FACET_DISPATCH(k_facet_vector, 16)
        FACET_DISPATCH(k_facet_vector, 21)
        FACET_DISPATCH(k_facet_vector, 25)
    default: return false;
    }
    return true;
}
#undef FACET_DISPATCH

void launch_coefficient_pts(int64_t np, const double *x, const double *y, const double *nx, const double *ny, int ninstr,
                            const int *code, const double *consts, double *out, hipStream_t s)
{
    if (np > 0) FEMD_LAUNCH(k_coefficient_pts, grid_for(np), TPB, 0, s, np, x, y, nx, ny, ninstr, code, consts, out);
}

void launch_coefficient(int nc, int nq, int nv, const double *geom, const double *qxi, const double *qeta, int ninstr,
                        const int *code, int nreg, const double *consts, const InputSet &in, double *out, hipStream_t s)
{
    const int64_t m = static_cast<int64_t>(nc) * nq;
    if (m <= 0) return;
    // a small register file lives in LDS (8 KiB per register row of the workgroup at most 16 rows),
    // a larger one in scratch memory; the arithmetic is the same
    if (nreg >= 1 && nreg <= 16)
        FEMD_LAUNCH(k_coefficient<true>, grid_for(m), TPB, size_t(nreg) * TPB * sizeof(double), s, nc, nq, nv, geom, qxi,
                    qeta, ninstr, code, nreg, consts, in, out);
    else
        FEMD_LAUNCH(k_coefficient<false>, grid_for(m), TPB, 0, s, nc, nq, nv, geom, qxi, qeta, ninstr, code, nreg, consts, in,
                    out);
}

#define FIELD_CASE(N, NV)                                                                                                  \
    case N:                                                                                                                \
        FEMD_LAUNCH((k_eval_field<N, NV>), grid_for(m), TPB, 0, s, nc, nq, cell_dof, u, tab, geom, qxi, qeta, code, out); \
        return true;

bool launch_eval_field(int nv, int nloc, int nc, int nq, const int *cell_dof, const double *u, const double *tab,
                       const double *geom, const double *qxi, const double *qeta, int code, double *out, hipStream_t s)
{
    const int64_t m = static_cast<int64_t>(nc) * nq;
    if (m <= 0) return true;
    if (nv == 4)
        switch (nloc)
        {
            FIELD_CASE(4, 4) FIELD_CASE(9, 4) FIELD_CASE(16, 4) FIELD_CASE(25, 4)
        default: return false;
        }
    switch (nloc)
    {
        FIELD_CASE(3, 3) FIELD_CASE(6, 3) FIELD_CASE(10, 3) FIELD_CASE(15, 3) FIELD_CASE(21, 3)
    default: return false;
    }
}
#undef FIELD_CASE

int nchunks(int64_t n) { return static_cast<int>((n + CHUNK - 1) / CHUNK); }

void launch_dot_parts(int64_t n, const double *x, const double *y, double *parts, const int *stop, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_dot_parts, nchunks(n), TPB, 0, s, n, x, y, parts, stop);
}

void launch_sum_parts(int np, const double *parts, double *out, hipStream_t s)
{
    FEMD_LAUNCH(k_sum_parts, 1, TPB, 0, s, np, parts, out);
}

void launch_sub(int64_t n, const double *b, const double *a, double *r, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_sub, grid_for(n), TPB, 0, s, n, b, a, r);
}
void launch_mul(int64_t n, const double *a, const double *b, double *c, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_mul, grid_for(n), TPB, 0, s, n, a, b, c);
}
void launch_copy(int64_t n, const double *a, double *b, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_copy, grid_for(n), TPB, 0, s, n, a, b);
}
void launch_fill(int64_t n, double v, double *a, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_fill, grid_for(n), TPB, 0, s, n, v, a);
}
void launch_gather_perm(int n, const int *perm, const double *src, double *dst, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_gather_perm, grid_for(n), TPB, 0, s, n, perm, src, dst);
}
void launch_scatter_perm(int n, const int *perm, const double *src, double *dst, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_scatter_perm, grid_for(n), TPB, 0, s, n, perm, src, dst);
}
void launch_take(int64_t n, const int *idx, const double *src, double *dst, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_take, grid_for(n), TPB, 0, s, n, idx, src, dst);
}
void launch_to_float(int64_t n, const double *a, float *b, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_to_float, grid_for(n), TPB, 0, s, n, a, b);
}
void launch_diag_positions(int n, const int *ip, const int *ix, int *diag, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_diag_positions, grid_for(n), TPB, 0, s, n, ip, ix, diag);
}
void launch_inv_diag(int n, const int *diag, const double *v, double *inv, int *bad_row, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_inv_diag, grid_for(n), TPB, 0, s, n, diag, v, inv, bad_row);
}
void launch_transpose_positions(int n, const int *ip, const int *ix, int *tpos, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_transpose_positions, grid_for(int64_t(n) * TP_LANES), TPB, 0, s, n, ip, ix, tpos);
}
void launch_asymmetry(int64_t nnz, const int *tpos, const double *v, double *out2, hipStream_t s)
{
    if (nnz > 0) FEMD_LAUNCH(k_asymmetry, std::min(grid_for(nnz), 4096), TPB, 0, s, nnz, tpos, v, out2);
}

void launch_cg_alpha(const double *parts, int np, CGState *st, hipStream_t s) { FEMD_LAUNCH(k_cg_alpha, 1, TPB, 0, s, parts, np, st); }
void launch_cg_update(int64_t n, double *x, double *r, const double *p, const double *Ap, const double *inv,
                      double *parts_rr, double *parts_rz, CGState *st, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_cg_update, nchunks(n), TPB, 0, s, n, x, r, p, Ap, inv, parts_rr, parts_rz, st);
}
void launch_cg_beta(const double *parts_rr, const double *parts_rz, int np, int mode, CGState *st, hipStream_t s)
{
    FEMD_LAUNCH(k_cg_beta, 1, TPB, 0, s, parts_rr, parts_rz, np, mode, st);
}
void launch_cg_rz(const double *parts, int np, CGState *st, hipStream_t s) { FEMD_LAUNCH(k_cg_rz, 1, TPB, 0, s, parts, np, st); }
void launch_cg_pupdate(int64_t n, double *p, const double *r, const double *inv, const CGState *st, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_cg_pupdate, grid_for(n), TPB, 0, s, n, p, r, inv, st);
}

void launch_ilu0_factor(int nlev, const int *lptr, const int *lrows, const int *ip, const int *ix, const int *diag,
                        double *v, int *bad_row, hipStream_t s)
{
    FEMD_LAUNCH(k_ilu0_factor, 1, WG_SWEEP, 0, s, nlev, lptr, lrows, ip, ix, diag, v, bad_row);
}
bool ilu0_levels_fit(int n, const int *ip, const int *ix, const int *diag)
{
    for (int i = 0; i < n; ++i)
    {
        if (ip[i + 1] - ip[i] > ILU_ROW) return false;
        int64_t upd = 0;
        for (int p = ip[i]; p < diag[i]; ++p) upd += ip[ix[p] + 1] - diag[ix[p]] - 1;
        if (upd > ILU_UPD) return false;
    }
    return true;
}
void launch_ilu0_levels(int nlev, const int *lptr_host, const int *lrows, const int *ip, const int *ix, const int *diag,
                        double *v, int *bad_row, hipStream_t s)
{
    for (int l = 0; l < nlev; ++l)
    {
        const int m = lptr_host[l + 1] - lptr_host[l];
        if (m > 0) FEMD_LAUNCH(k_ilu0_level, (m + ILU_WAVES - 1) / ILU_WAVES, 64 * ILU_WAVES, 0, s, m, lrows + lptr_host[l], ip, ix, diag, v, bad_row);
    }
}
void launch_min_pivot(int n, const int *diag, const double *v, double *out, hipStream_t s)
{
    FEMD_LAUNCH(k_min_pivot, 1, 1024, 0, s, n, diag, v, out);
}
void launch_ilu0_apply(int n, int nlevL, const int *lptrL, const int *lrowsL, int nlevU, const int *lptrU,
                       const int *lrowsU, const int *ip, const int *ix, const int *diag, const double *v,
                       const double *r, double *z, const int *stop, hipStream_t s)
{
    (void)n;
    FEMD_LAUNCH(k_ilu0_apply, 1, WG_SWEEP, 0, s, nlevL, lptrL, lrowsL, nlevU, lptrU, lrowsU, ip, ix, diag, v, r, z, stop);
}
void launch_ssor_apply(int n, int nlevL, const int *lptrL, const int *lrowsL, int nlevU, const int *lptrU,
                       const int *lrowsU, const int *ip, const int *ix, const int *diag, const double *v, double omega,
                       const double *r, double *z, const int *stop, hipStream_t s)
{
    FEMD_LAUNCH(k_ssor_apply, 1, WG_SWEEP, 0, s, n, nlevL, lptrL, lrowsL, nlevU, lptrU, lrowsU, ip, ix, diag, v, omega, r, z, stop);
}

int fast_blocks(int64_t n)
{
    const int64_t g = (n + 4 * TPB - 1) / (4 * TPB);
    return static_cast<int>(g < 1 ? 1 : (g > FAST_MAX_BLOCKS ? FAST_MAX_BLOCKS : g));
}
void launch_fdot(int64_t n, const double *x, const double *y, double *parts, hipStream_t s)
{
    FEMD_LAUNCH(k_fdot, fast_blocks(n), TPB, 0, s, n, x, y, parts);
}
void launch_fsum(int np, const double *parts, double *out, hipStream_t s) { FEMD_LAUNCH(k_fsum, 1, TPB, 0, s, np, parts, out); }
void launch_axpy_dev(int64_t n, const double *alpha, double sign, const double *x, double *y, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_axpy_dev, grid_for(n), TPB, 0, s, n, alpha, sign, x, y);
}
void launch_scale_dev(int64_t n, const double *num, const double *den, double *x, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_scale_dev, grid_for(n), TPB, 0, s, n, num, den, x);
}
void launch_axpy(int64_t n, double a, const double *x, double *y, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_axpy, grid_for(n), TPB, 0, s, n, a, x, y);
}
void launch_scale(int64_t n, double a, double *x, hipStream_t s)
{
    if (n > 0) FEMD_LAUNCH(k_scale, grid_for(n), TPB, 0, s, n, a, x);
}
void launch_lincomb(int64_t n, int k, const double *y, const double *const *V, double *x, hipStream_t s)
{
    for (int j0 = 0; j0 < k; j0 += 64)
    {
        VecPtrs P{};
        const int kk = k - j0 < 64 ? k - j0 : 64;
        for (int j = 0; j < kk; ++j) P.p[j] = V[j0 + j];
        if (n > 0) FEMD_LAUNCH(k_lincomb, grid_for(n), TPB, 0, s, n, kk, y + j0, P, x);
    }
}


} // namespace femd_amd
