// cg32.hip -- the FP32 inner CG: the scaled FP32 copy of A, and the kernels and host driver of
// one inner solve (the product fused into the CG, k_scg_spmv, is in spmv.hip).
//
// Iteration k:
//   k_scg_spmv    w = (S A S) r, partials of r.r and w.r per workgroup (spmv.hip)
//   k_scg_scalars gamma, delta from those partials; stop when gamma <= tol^2 gamma_0, at
//                 max_iter, or on NaN / breakdown; alpha, beta
//   k_scg_update  p, s, x, r updated in place
// Every kernel returns at once when the state says stop, so a fixed sequence of launches (or a
// captured graph of them) is safe to run past convergence.
#include "cg32.hpp"

#include <cmath>

namespace femd_amd {
namespace {

constexpr int TB = 256;
constexpr int MAX_GRID = 1024;

inline int vec_grid(int64_t n) { return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(MAX_GRID, (n + TB - 1) / TB))); }

__global__ __launch_bounds__(TB) void k_scale_of_diag(int n, const int *diag, const double *v, double *sc, int *bad_row)
{
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int d = diag[i];
    const double a = d >= 0 ? v[d] : 0.0;
    if (!(a > 0.0))
    {
        atomicMin(bad_row, i);
        sc[i] = 0.0;
    }
    else sc[i] = 1.0 / std::sqrt(a);
}

// out[p] = v[p] s_i s_j for the entries of row i (one lane per row: a set-up pass).
__global__ __launch_bounds__(TB) void k_scaled_values(int n, const int *ip, const int *ix, const double *v, const double *sc,
                                                      float *out)
{
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const double si = sc[i];
    for (int p = ip[i]; p < ip[i + 1]; ++p) out[p] = static_cast<float>(v[p] * si * sc[ix[p]]);
}

// r = S r64 / sqrt(rr64), y = p = s = 0, and the state of a new solve.
__global__ __launch_bounds__(TB) void k_scg_init(int64_t n, const double *r64, const double *rr64, const double *sc, float *r,
                                                 float *x32, float *p, float *sv, SCGState *st, double tol2, int max_iter)
{
    const double rr = rr64[0];
    const double scale = rr > 0.0 ? 1.0 / std::sqrt(rr) : 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * TB)
    {
        r[i] = static_cast<float>(sc[i] * r64[i] * scale);
        x32[i] = 0.0f;
        p[i] = 0.0f;
        sv[i] = 0.0f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        st->gam_prev = 0.0;
        st->alp_prev = 0.0;
        st->alpha = st->beta = 0.0;
        st->thr = 0.0;
        st->tol2 = tol2;
        st->rr = 0.0;
        st->rho = rr > 0.0 ? std::sqrt(rr) : 0.0;
        st->it = 0;
        st->max_iter = max_iter;
        st->breakdown = 0;
        st->fresh = 1;
        st->stop = rr > 0.0 ? 0 : 1;
    }
}

// Iteration k after the product w = A r (one workgroup): gamma = r.r and delta = w.r from the
// product's partials in a fixed order, the stopping test, then (Chronopoulos-Gear)
//   beta = gamma / gamma_{k-1},  alpha = gamma / (delta - beta gamma / alpha_{k-1})   (k > 0)
//   beta = 0,                    alpha = gamma / delta                                (k = 0).
constexpr int TS = 1024;
__global__ __launch_bounds__(TS) void k_scg_scalars(const double *parts, int nparts, SCGState *st, int k)
{
    if (st->stop) return;
    __shared__ double red[4 * (TS / 64)];
    // the quadruples (r.r, w.r, w.p, p.s) in a fixed order: lane-strided, wave shuffles, waves
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    const double2 *pp = reinterpret_cast<const double2 *>(parts);
    for (int i = threadIdx.x; i < nparts; i += TS)
    {
        const double2 t0 = pp[2 * i], t1 = pp[2 * i + 1];
        q[0] += t0.x;
        q[1] += t0.y;
        q[2] += t1.x;
        q[3] += t1.y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] += __shfl_xor(q[j], off, 64);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[4 * (threadIdx.x >> 6) + j] = q[j];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int j = 0; j < 4; ++j) q[j] = 0.0;
    for (int wv = 0; wv < TS / 64; ++wv)
        for (int j = 0; j < 4; ++j) q[j] += red[4 * wv + j];
    const double gam = q[0], del = q[1], wp = q[2], ps = q[3];
    const double thr = st->fresh ? st->tol2 * gam : st->thr;
    const bool nan = !(gam == gam) || !(del == del);
    if (nan || gam <= thr || k >= st->max_iter)
    {
        st->stop = 1;
        st->breakdown = nan ? 1 : 0;
        st->rr = gam;
        st->it = k;
        return;
    }
    double beta = 0.0, den = del;
    if (k > 0 && st->fresh)
    {
        // r was replaced by the true residual: the beta that makes p' = r + beta p conjugate to p,
        // -(r.A p)/(p.A p) = -(w.p)/(p.s), and p'.A p' = w.r + 2 beta w.p + beta^2 p.s exactly
        beta = ps > 0.0 ? -wp / ps : 0.0;
        den = del + 2.0 * beta * wp + beta * beta * ps;
    }
    else if (k > 0)
    {
        // r from the recurrence (Chronopoulos-Gear)
        beta = gam / st->gam_prev;
        den = del - beta * gam / st->alp_prev;
    }
    if (!(den > 0.0))
    {
        st->stop = 1;
        st->breakdown = 1;
        st->rr = gam;
        st->it = k;
        return;
    }
    const double alpha = gam / den;
    st->alpha = alpha;
    st->beta = beta;
    st->gam_prev = gam;
    st->alp_prev = alpha;
    if (st->fresh) st->thr = thr;
    st->fresh = 0;
    st->it = k + 1;
}

// p = r + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s.
__global__ __launch_bounds__(TB) void k_scg_update(int64_t n, float *x32, float *r, const float *w, float *p, float *sv,
                                                   const SCGState *st)
{
    if (st->stop) return;
    const float a = static_cast<float>(st->alpha), bt = static_cast<float>(st->beta);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * TB)
    {
        const float pi = fmaf(bt, p[i], r[i]);
        const float si = fmaf(bt, sv[i], w[i]);
        p[i] = pi;
        sv[i] = si;
        x32[i] = fmaf(a, pi, x32[i]);
        r[i] = fmaf(-a, si, r[i]);
    }
}

__global__ void k_scg_finish(int64_t n, const double *sc, const float *x32, const SCGState *st, double *x64)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rho = st->rho;
    if (rho > 0.0) x64[i] += rho * (sc[i] * static_cast<double>(x32[i]));
}

// A new FP64 residual r64 with norm^2 rr64[0]: r = S r64 / rho_new, y = 0, p and s times
// rho / rho_new; the state keeps its direction (previous r.r rescaled) and starts a segment.
__global__ __launch_bounds__(256) void k_scg_rescale(int64_t n, const double *r64, const double *rr64, const double *sc,
                                                     float *r, float *x32, float *p, float *sv, SCGState *st)
{
    const double rr = rr64[0];
    const double rho_new = rr > 0.0 ? std::sqrt(rr) : 0.0;
    const double rho = st->rho;
    const double scale = rho_new > 0.0 ? 1.0 / rho_new : 0.0;
    const float f = static_cast<float>(rho_new > 0.0 ? rho / rho_new : 0.0);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256)
    {
        r[i] = static_cast<float>(sc[i] * r64[i] * scale);
        x32[i] = 0.0f;
        p[i] *= f;
        sv[i] *= f;
    }
    __syncthreads();
}

__global__ void k_scg_rescale_state(const double *rr64, SCGState *st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double rr = rr64[0];
    const double rho_new = rr > 0.0 ? std::sqrt(rr) : 0.0;
    const double f = rho_new > 0.0 ? st->rho / rho_new : 0.0;
    st->gam_prev *= f * f;
    st->rho = rho_new;
    st->fresh = 1;
    st->breakdown = 0;
    st->stop = rho_new > 0.0 ? 0 : 1;
}

} // namespace

// S A S in FP32 and S, cached on the matrix until its values change.
static void scaled_values(femd_csr_s *A, hipStream_t s)
{
    if (A->v32s_valid) return;
    DevPattern &P = *A->pat;
    if (P.n != P.m) fail(FEMD_EINVAL, "inner CG: the matrix must be square");
    const int n = P.n;
    const int *diag = diag_of(P, s);
    A->scale.ensure(sizeof(double) * std::max(n, 1));
    A->v32s.ensure(sizeof(float) * std::max<int64_t>(P.nnz, 1));
    DevBuf bad(sizeof(int));
    const int big = 0x7fffffff;
    HIPC(hipMemcpyAsync(bad.p, &big, sizeof(int), hipMemcpyHostToDevice, s));
    if (n > 0)
    {
        hipLaunchKernelGGL(k_scale_of_diag, dim3((n + TB - 1) / TB), dim3(TB), 0, s, n, diag, A->v.as<double>(),
                           A->scale.as<double>(), bad.as<int>());
        hipLaunchKernelGGL(k_scaled_values, dim3((n + TB - 1) / TB), dim3(TB), 0, s, n, P.ip.as<int>(), P.ix.as<int>(),
                           A->v.as<double>(), A->scale.as<double>(), A->v32s.as<float>());
    }
    int badrow = big;
    HIPC(hipMemcpyAsync(&badrow, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (badrow != big)
        fail(FEMD_ENOTSPD, "inner CG: diagonal entry of row " + std::to_string(badrow) + " is not positive");
    A->v32s_valid = true;
}

void ScaledCG::setup(femd_csr_s *A, hipStream_t s)
{
    scaled_values(A, s);
    P = A->pat.get();
    n = P->n;
    v = A->v32s.as<float>();
    sc = A->scale.as<double>();
    plan = &spmv_plan(*P, FEMD_FP32, v, s);
    const size_t N = size_t(std::max(n, 1));
    r.ensure(4 * N);
    w.ensure(4 * N);
    p.ensure(4 * N);
    sv.ensure(4 * N);
    x32.ensure(4 * N);
    parts.ensure(32 * size_t(std::max(plan->ntiles, 1)));
    if (!st.p)
    {
        st.ensure(sizeof(SCGState));
        SCGState z{};
        HIPC(hipMemcpyAsync(st.p, &z, sizeof(z), hipMemcpyHostToDevice, s));
        HIPC(hipStreamSynchronize(s));
    }
}

void ScaledCG::init(const double *r64, const double *rr64, double inner_tol, int max_iter, hipStream_t s)
{
    hipLaunchKernelGGL(k_scg_init, dim3(vec_grid(n)), dim3(TB), 0, s, static_cast<int64_t>(n), r64, rr64, sc, r.as<float>(),
                       x32.as<float>(), p.as<float>(), sv.as<float>(), st.as<SCGState>(), inner_tol * inner_tol, max_iter);
}

// Sampled event timing of the fused product (every 8th iteration of solves outside graph capture),
// so a program can read the kernel's average duration in the middle of real solves.
namespace {
struct Sample {
    hipEvent_t e0, e1;
    int k;
    bool valid;
};
struct InnerTimer {
    std::mutex mu;
    bool on = false;
    std::vector<Sample> samples;
    size_t used = 0;
};
InnerTimer g_timer;

bool timer_sample(hipStream_t s, int k, hipEvent_t *e0, hipEvent_t *e1)
{
    std::lock_guard<std::mutex> lk(g_timer.mu);
    if (!g_timer.on) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    if (g_timer.used == g_timer.samples.size())
    {
        Sample x{nullptr, nullptr, 0, false};
        HIPC(hipEventCreate(&x.e0));
        HIPC(hipEventCreate(&x.e1));
        g_timer.samples.push_back(x);
    }
    Sample &x = g_timer.samples[g_timer.used++];
    x.k = k;
    x.valid = true;
    *e0 = x.e0;
    *e1 = x.e1;
    return true;
}

size_t timer_mark()
{
    std::lock_guard<std::mutex> lk(g_timer.mu);
    return g_timer.used;
}

// Samples from `mark` on whose iteration came after the solve stopped timed a no-op launch.
void timer_drop_after(size_t mark, int iterations)
{
    std::lock_guard<std::mutex> lk(g_timer.mu);
    for (size_t i = mark; i < g_timer.used; ++i)
        if (g_timer.samples[i].k >= iterations) g_timer.samples[i].valid = false;
}
} // namespace

void ScaledCG::iterate(int k0, int count, hipStream_t s)
{
    for (int k = k0; k < k0 + count; ++k)
    {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if ((k & 7) != 0 || !timer_sample(s, k, &e0, &e1)) e0 = e1 = nullptr;
        const int np = scg_spmv(*plan, *P, v, r.as<float>(), p.as<float>(), sv.as<float>(), w.as<float>(), parts.as<double>(),
                                st.as<SCGState>(), k, s, e0, e1);
        hipLaunchKernelGGL(k_scg_scalars, dim3(1), dim3(TS), 0, s, parts.as<double>(), np, st.as<SCGState>(), k);
        hipLaunchKernelGGL(k_scg_update, dim3(vec_grid(n)), dim3(TB), 0, s, static_cast<int64_t>(n), x32.as<float>(),
                           r.as<float>(), w.as<float>(), p.as<float>(), sv.as<float>(), st.as<SCGState>());
    }
}

void ScaledCG::finish(double *x64, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_scg_finish, dim3((n + TB - 1) / TB), dim3(TB), 0, s, static_cast<int64_t>(n), sc, x32.as<float>(),
                           st.as<SCGState>(), x64);
}

void ScaledCG::rescale(const double *r64, const double *rr64, hipStream_t s)
{
    hipLaunchKernelGGL(k_scg_rescale, dim3(vec_grid(n)), dim3(256), 0, s, static_cast<int64_t>(n), r64, rr64, sc, r.as<float>(),
                       x32.as<float>(), p.as<float>(), sv.as<float>(), st.as<SCGState>());
    hipLaunchKernelGGL(k_scg_rescale_state, dim3(1), dim3(64), 0, s, rr64, st.as<SCGState>());
}

SCGState ScaledCG::resume(int *k, hipStream_t s)
{
    Pinned h(sizeof(SCGState));
    const size_t mark = timer_mark();
    SCGState out{};
    int batch = 8;
    for (;;)
    {
        HIPC(hipMemcpyAsync(h.p, st.p, sizeof(SCGState), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        std::memcpy(&out, h.p, sizeof(out));
        if (out.stop) break;
        iterate(*k, batch, s);
        *k += batch;
        batch = std::min(64, 2 * batch);
    }
    *k = out.it;
    timer_drop_after(mark, out.it);
    return out;
}

SCGState ScaledCG::run(const double *r64, const double *rr64, double inner_tol, int max_iter, hipStream_t s)
{
    init(r64, rr64, inner_tol, max_iter, s);
    int k = 0;
    return resume(&k, s);
}

ScaledCG::Result ScaledCG::reliable(femd_csr_s *A, const double *b, double *x, double tol, double inner_tol, int max_updates,
                                    int max_iter, hipStream_t s)
{
    Result res;
    DevBuf r64(8 * size_t(std::max(n, 1))), t64(8 * size_t(std::max(n, 1))), parts64(8 * 2048 + 64), scal(64);
    const int nbk = fast_blocks(n);
    auto residual = [&]() {
        // FP64 r = b - A x and its squared norm in scal[0]
        spmv_run64(spmv_plan(*A->pat, FEMD_FP64, A->v.p, s), *A->pat, A->v.as<double>(), x, t64.as<double>(), nullptr, s);
        launch_sub(n, b, t64.as<double>(), r64.as<double>(), s);
        launch_fdot(n, r64.as<double>(), r64.as<double>(), parts64.as<double>(), s);
        launch_fsum(nbk, parts64.as<double>(), scal.as<double>(), s);
        return std::sqrt(read_scalar(scal.as<double>(), s));
    };
    launch_fdot(n, b, b, parts64.as<double>(), s);
    launch_fsum(nbk, parts64.as<double>(), scal.as<double>(), s);
    const double nb = std::sqrt(read_scalar(scal.as<double>(), s));
    if (nb == 0.0)
    {
        launch_fill(n, 0.0, x, s);
        HIPC(hipStreamSynchronize(s));
        res.converged = 1;
        return res;
    }
    double rn = residual();
    res.residual = rn / nb;
    if (res.residual <= tol) { res.converged = 1; return res; }
    init(r64.as<double>(), scal.as<double>(), inner_tol, max_iter, s);
    int k = 0, done = 0;
    for (;;)
    {
        const SCGState out = resume(&k, s);
        finish(x, s);
        ++res.updates;
        rn = residual();
        res.residual = rn / nb;
        res.iterations = done + k;
        if (res.residual <= tol) { res.converged = 1; break; }
        if (res.iterations >= max_iter || res.updates >= max_updates) break;
        if (out.breakdown)
        {
            // a non-finite or non-positive scalar: start over from the true residual
            res.breakdown += 1;
            done += k;
            init(r64.as<double>(), scal.as<double>(), inner_tol, max_iter - done, s);
            k = 0;
            continue;
        }
        rescale(r64.as<double>(), scal.as<double>(), s);
    }
    return res;
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
            if (!x.valid) continue;
            HIPC(hipEventSynchronize(x.e1));
            float ms = 0.0f;
            const hipError_t e = hipEventElapsedTime(&ms, x.e0, x.e1);
            if (e == hipErrorNotReady) continue;
            hip_check(e, "hipEventElapsedTime");
            t += ms;
            ++k;
        }
        if (launches) *launches = k;
        if (total_ms) *total_ms = t;
    });
}

} // extern "C"
