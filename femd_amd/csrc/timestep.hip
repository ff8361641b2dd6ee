// timestep.hip -- linear time stepping with a fixed matrix pair, each step one replay of a
// captured hipGraph (the time-stepping pattern; the reference writes this loop in Python, e.g.
// example13.ipynb: b = M u, then a solve with A, every step).
//
// Backward Euler for M u' + K u = 0 is  A u_{n+1} = M u_n  with A = M + dt K.  One step:
//   b = M u (FP64, the reference's row order), bb = b.b;
//   `rounds` times: r = b - A u (FP64), rr = r.r; unless rr <= tol^2 bb, a segment of at most
//   max_inner iterations of the scaled Jacobi-CG (cg_engine.hpp, one system, FP32 or FP64) on
//   A d = r / sqrt(rr) from d = 0, until its residual has dropped by inner_tol or its estimate of
//   ||b - A u|| meets tol ||b||, and u += sqrt(rr) d (FP64).  The first round of a step starts
//   with no search direction (the step starts from u_n: a warm start); the later rounds keep the
//   direction, conjugate to the next one -- reliable updates of one CG run, not restarts;
//   the step's inner iterations are added to a per-step counter on the device.
// Every scalar lives on the device and every kernel of a finished solve returns at once, so the
// whole step is one fixed launch sequence: captured once, replayed nsteps times back to back, and
// the host reads the counters at the end.
#include "cg_engine.hpp"
#include "kernels.hpp"

#include <cmath>

namespace femd_amd {
namespace {

struct StepScalars {
    double bb;          // ||b||^2 of the current step
    double rr;          // ||b - A u||^2 of the current round (0 when the round is skipped)
    int counter;        // steps done
    int total_inner;
    int seen;           // the engine's iteration total at the last record
    int step_inner;     // the current step's inner iterations
    double rr_end;      // ||b - A u||^2 once the step's rounds are done
    double worst;       // the largest ||b - A u||^2 / ||b||^2 of any step so far
    int failed;         // steps the verdict of krylov/cg.hpp:222-237 does not accept
    int pad;
};

__global__ void k_gate(StepScalars *sc, double tol2)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && sc->rr <= tol2 * sc->bb) sc->rr = 0.0;
}

// the norm the inner solve scales by (0: the round is skipped) and its target (tol ||b|| / rho)^2
__global__ void k_rho(const StepScalars *sc, double tol2, double *rho, double *tgt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *rho = sc->rr > 0.0 ? sqrt(sc->rr) : 0.0;
    *tgt = sc->rr > 0.0 ? tol2 * sc->bb / sc->rr : 0.0;
}

__global__ void k_record(const CGSys *st, StepScalars *sc, int *iters, int nmax)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int it = st->total - sc->seen;
    sc->seen = st->total;
    if (iters && sc->counter < nmax) iters[sc->counter] += it;
    sc->total_inner += it;
    sc->step_inner += it;
}

// The step's verdict, the reference's for one CG solve (krylov/cg.hpp:222-237): its true relative
// residual within tol (1 + 1e-6) + 1e-15, or within 10 tol when its rounds stopped before their
// iteration budget ran out (an FP64 residual near its attainable accuracy).  Then the next step.
__global__ void k_next_step(StepScalars *sc, double tol, int budget)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double rel2 = sc->bb > 0.0 ? sc->rr_end / sc->bb : sc->rr_end;
    if (!(rel2 <= sc->worst)) sc->worst = rel2;        // NaN sticks
    const double rel = sqrt(rel2);
    const bool ok = rel <= tol * (1.0 + 1e-6) + 1e-15 || (rel <= 10.0 * tol && sc->step_inner < budget);
    if (!ok) sc->failed += 1;
    sc->step_inner = 0;
    sc->counter += 1;
}

struct Stepper {
    femd_csr_s *M = nullptr, *A = nullptr;
    int n = 0, rounds = 1, max_inner = 1, nsteps = 0, precision = FEMD_FP32;
    double tol = 0.0, inner_tol = 0.0;
    const SpmvPlan *pM = nullptr, *pA = nullptr;
    CGEngine cg;
    DevBuf b, t, r, parts, sc, iters, rho, tgt, sel;     // sel: {2, 1}, restart and resume
    double *u = nullptr;

    StepScalars *scal() const { return sc.as<StepScalars>(); }

    void norm2(const double *x, double *out, hipStream_t s)
    {
        launch_fdot(n, x, x, parts.as<double>(), s);
        launch_fsum(fast_blocks(n), parts.as<double>(), out, s);
    }

    void step(hipStream_t s)
    {
        spmv_run64(*pM, *M->pat, M->v.as<double>(), u, b.as<double>(), nullptr, s);
        norm2(b.as<double>(), &scal()->bb, s);
        for (int k = 0; k < rounds; ++k)
        {
            spmv_run64(*pA, *A->pat, A->v.as<double>(), u, t.as<double>(), nullptr, s);
            launch_sub(n, b.as<double>(), t.as<double>(), r.as<double>(), s);
            norm2(r.as<double>(), &scal()->rr, s);
            hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, s, scal(), tol * tol);
            hipLaunchKernelGGL(k_rho, dim3(1), dim3(64), 0, s, scal(), tol * tol, rho.as<double>(), tgt.as<double>());
            cg.start_device(r.as<double>(), rho.as<double>(), tgt.as<double>(), sel.as<int>() + (k == 0 ? 0 : 1), inner_tol,
                            max_inner, s);
            cg.iterate(max_inner, s);
            cg.finish_device(sel.as<int>(), u, s);
            hipLaunchKernelGGL(k_record, dim3(1), dim3(64), 0, s, cg.st.as<CGSys>(), scal(), iters.as<int>(), nsteps);
        }
        // the step's own FP64 residual, so that every step (not just the last) is checked
        spmv_run64(*pA, *A->pat, A->v.as<double>(), u, t.as<double>(), nullptr, s);
        launch_sub(n, b.as<double>(), t.as<double>(), r.as<double>(), s);
        norm2(r.as<double>(), &scal()->rr_end, s);
        hipLaunchKernelGGL(k_next_step, dim3(1), dim3(64), 0, s, scal(), tol, rounds * max_inner);
    }
};

void heat_run(femd_csr_s *M, femd_csr_s *A, double *u, int nsteps, double tol, int rounds, double inner_tol, int max_inner,
              int precision, int use_graph, int *iters_out, int *failed_out, femd_solve_info *info, hipStream_t s)
{
    const DevPattern &PM = *M->pat, &PA = *A->pat;
    const int n = PA.n;
    if (PA.m != n || PM.n != n || PM.m != n) fail(FEMD_EINVAL, "time stepping: M and A must be square and of one size");
    if (nsteps < 0 || rounds < 1 || max_inner < 1 || !(tol >= 0.0) || !(inner_tol > 0.0 && inner_tol < 1.0))
        fail(FEMD_EINVAL, "time stepping: need nsteps >= 0, rounds >= 1, max_inner >= 1, tol >= 0, 0 < inner_tol < 1");
    if (precision != FEMD_FP32 && precision != FEMD_FP64) fail(FEMD_EINVAL, "time stepping: precision is FEMD_FP32 or FEMD_FP64");
    std::memset(info, 0, sizeof(*info));
    if (iters_out && nsteps > 0) std::memset(iters_out, 0, sizeof(int) * size_t(nsteps));
    if (failed_out) *failed_out = 0;
    if (n == 0) { info->converged = 1; return; }
    Stepper T;
    T.M = M; T.A = A; T.n = n; T.rounds = rounds; T.max_inner = max_inner; T.nsteps = nsteps;
    T.tol = tol; T.inner_tol = inner_tol; T.u = u; T.precision = precision;
    // everything that tunes, allocates or synchronizes happens before the capture
    T.pM = &spmv_plan(*M->pat, FEMD_FP64, M->v.p, s);
    T.pA = &spmv_plan(*A->pat, FEMD_FP64, A->v.p, s);
    T.cg.setup(A, precision, s);
    const size_t N = size_t(n);
    T.b = DevBuf(8 * N); T.t = DevBuf(8 * N); T.r = DevBuf(8 * N);
    T.parts = DevBuf(8 * 2048 + 64);
    T.sc = DevBuf(sizeof(StepScalars));
    T.rho = DevBuf(sizeof(double));
    T.tgt = DevBuf(sizeof(double));
    T.sel = DevBuf(2 * sizeof(int));
    const int modes[2] = {2, 1};
    HIPC(hipMemcpyAsync(T.sel.p, modes, sizeof(modes), hipMemcpyHostToDevice, s));
    T.iters = DevBuf(sizeof(int) * size_t(std::max(nsteps, 1)));
    const StepScalars z{};
    HIPC(hipMemcpyAsync(T.sc.p, &z, sizeof(z), hipMemcpyHostToDevice, s));
    const std::vector<int> zi(size_t(std::max(nsteps, 1)), 0);
    HIPC(hipMemcpyAsync(T.iters.p, zi.data(), sizeof(int) * zi.size(), hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
    if (nsteps > 0)
    {
        if (use_graph)
        {
            hipGraph_t g = nullptr;
            hipGraphExec_t ex = nullptr;
            HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try { T.step(s); }
            catch (...) { (void)hipStreamEndCapture(s, &g); if (g) (void)hipGraphDestroy(g); throw; }
            HIPC(hipStreamEndCapture(s, &g));
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (e != hipSuccess) { (void)hipGraphDestroy(g); hip_check(e, "hipGraphInstantiate"); }
            hipError_t le = hipSuccess;
            for (int k = 0; k < nsteps && le == hipSuccess; ++k) le = hipGraphLaunch(ex, s);
            const hipError_t se = hipStreamSynchronize(s);
            (void)hipGraphExecDestroy(ex);
            (void)hipGraphDestroy(g);
            hip_check(le, "hipGraphLaunch");
            hip_check(se, "hipStreamSynchronize");
        }
        else
        {
            for (int k = 0; k < nsteps; ++k) T.step(s);
            HIPC(hipStreamSynchronize(s));
        }
    }
    // every step's relative residual ||b - A u|| / ||b|| after its rounds: the worst of them
    double res = 0.0;
    StepScalars hs{};
    if (nsteps > 0)
    {
        HIPC(hipMemcpyAsync(&hs, T.sc.p, sizeof(hs), hipMemcpyDeviceToHost, s));
        if (iters_out) HIPC(hipMemcpyAsync(iters_out, T.iters.p, sizeof(int) * size_t(nsteps), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        res = std::sqrt(hs.worst);
    }
    info->iterations = hs.total_inner;
    info->breakdown = hs.counter;
    info->residual = res;
    info->precond_residual = res;
    info->converged = (hs.counter == nsteps && hs.failed == 0) ? 1 : 0;
    if (failed_out) *failed_out = hs.failed;
}

} // namespace
} // namespace femd_amd

using namespace femd_amd;

extern "C" {

int femd_heat_run_dev(femd_csr M, femd_csr A, double *u_dev, int nsteps, double tol, int rounds, double inner_tol,
                      int max_inner, int precision, int use_graph, int *iters, int *failed, femd_solve_info *info,
                      femd_stream st)
{
    return api([&] {
        if (!info) fail(FEMD_EINVAL, "time stepping: info is NULL");
        heat_run(check_csr(M), check_csr(A), u_dev, nsteps, tol, rounds, inner_tol, max_inner, precision, use_graph, iters,
                 failed, info, S(st));
    });
}

int femd_heat_run(femd_csr M, femd_csr A, double *u, int nsteps, double tol, int rounds, double inner_tol, int max_inner,
                  int precision, int use_graph, int *iters, int *failed, femd_solve_info *info)
{
    return api([&] {
        if (!info) fail(FEMD_EINVAL, "time stepping: info is NULL");
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        const int n = a->pat->n;
        DevBuf ud = upload(u, size_t(n), s);
        heat_run(check_csr(M), a, ud.as<double>(), nsteps, tol, rounds, inner_tol, max_inner, precision, use_graph, iters,
                 failed, info, s);
        if (n) HIPC(hipMemcpyAsync(u, ud.p, 8 * size_t(n), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

} // extern "C"
