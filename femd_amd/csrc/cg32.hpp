// cg32.hpp -- the FP32 inner CG shared by the refinement solver, the time stepper and the serving
// loop: its device state, the block-redundant reduction it is built on, and the host driver.
//
// The inner solve is Jacobi-preconditioned CG written as plain CG on the symmetrically scaled
// system  (S A S) y = S r,  S = diag(a_ii)^(-1/2),  d = S y  (the same iterates as PCG with
// M = diag(A) in exact arithmetic), in the Chronopoulos-Gear arrangement (one reduction phase per
// iteration: both dot products come out of the product kernel).  Vectors are FP32, scalars and
// reductions FP64.
//
// An iteration is three kernels: the product w = A r with the partials of r.r and w.r per
// workgroup; one workgroup that adds those partials in a fixed order and sets alpha, beta and
// the stopping flag; the vector update.  No atomics or fences, and the result does not depend on
// timing.  For mixed-precision solves with reliable updates the FP32 residual is replaced by
// the FP64 one from time to time; the first iteration after that takes p'.A p' from explicit dot
// products (w.p, p.s) instead of the recurrence, which assumes the residual came from it.
#pragma once
#include "internal.hpp"

namespace femd_amd {

struct SCGState {
    double gam_prev;    // r.r of the previous iteration
    double alp_prev;    // alpha of the previous iteration
    double alpha, beta; // of the current iteration, for the update kernel
    double thr;         // stop once r.r <= thr  (tol^2 times r.r at the start of the segment)
    double tol2;
    double rr;          // r.r when stopped
    double rho;         // the FP64 residual norm the FP32 system is scaled by
    int it, max_iter, stop, breakdown;
    int fresh;          // the next iteration starts a segment: it sets thr from its r.r
};

// Two FP64 partial sums carried together.
struct Dot2 {
    double a = 0.0, b = 0.0;
    __host__ __device__ Dot2 &operator+=(const Dot2 &o)
    {
        a += o.a;
        b += o.b;
        return *this;
    }
};

// The sum of a[0..n) in a fixed order (lane-strided, then wave shuffles, then waves in order),
// returned to every thread of the workgroup.  a has stride `stride` doubles.
template <int THREADS>
__device__ inline double wg_fixed_sum(const double *a, int n, double *red, int stride = 1)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += THREADS) s += a[static_cast<int64_t>(i) * stride];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < THREADS / 64; ++w) t += red[w];
    __syncthreads();
    return t;
}

// The product of the FP32 CG of iteration k (spmv.hip): w = (S A S) r and, per workgroup b, the
// partials of r.r, w.r, w.p, p.s over its own rows into parts[4 b .. 4 b + 3] (the last two only
// on the first iteration of a segment after the first, else zero).  Returns the number of
// workgroups.  With e0 / e1 not NULL the launch is bracketed by those events.
int scg_spmv(const SpmvPlan &p, const DevPattern &P, const float *v, const float *r, const float *pv, const float *sv, float *w,
             double *parts, const SCGState *st, int k, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

// Work space and launch sequence of one inner solve on one matrix.
struct ScaledCG {
    int n = 0;
    DevPattern *P = nullptr;
    const float *v = nullptr;            // S A S in FP32
    const double *sc = nullptr;          // S
    const SpmvPlan *plan = nullptr;
    DevBuf r, w, p, sv, x32, parts, st;        // sv: the CG vector s = A p

    // Scale values and work buffers for A (sets v, sc, plan).
    void setup(femd_csr_s *A, hipStream_t s);
    // Start a solve of (S A S) y = S r64 / sqrt(rr64[0]) from y = 0, to ||r_k|| <= inner_tol ||r_0||
    // or max_iter iterations; rr64[0] == 0 starts a solve that is already finished.
    void init(const double *r64, const double *rr64, double inner_tol, int max_iter, hipStream_t s);
    // Iterations k0 .. k0+count-1; no-ops once stopped.
    void iterate(int k0, int count, hipStream_t s);
    // x64 += rho S y, the solve's correction in FP64.
    void finish(double *x64, hipStream_t s);
    // Keep going from a new FP64 residual r64 (norm^2 rr64[0]) after finish: r = S r64 / rho_new,
    // y = 0, and p, s (and the previous r.r) rescaled by rho / rho_new, so the search direction
    // survives; the next iteration starts a segment to inner_tol of its own r.r.
    void rescale(const double *r64, const double *rr64, hipStream_t s);
    // init, then iterate in growing batches from index k0, polling the state, until it stops.
    SCGState run(const double *r64, const double *rr64, double inner_tol, int max_iter, hipStream_t s);
    SCGState resume(int *k, hipStream_t s);
    // Mixed-precision CG with reliable updates: one FP32 CG (never restarted) on the scaled
    // system; whenever its r.r has dropped by inner_tol^2, its correction goes into x (FP64), the
    // true residual b - A x is recomputed in FP64 and the FP32 vectors rescaled to it.  Stops when
    // ||b - A x|| <= tol ||b||, after max_updates updates or max_iter iterations.
    struct Result {
        int converged = 0, iterations = 0, updates = 0, breakdown = 0;
        double residual = 0.0;
    };
    Result reliable(femd_csr_s *A, const double *b, double *x, double tol, double inner_tol, int max_updates, int max_iter,
                    hipStream_t s);
};

} // namespace femd_amd
