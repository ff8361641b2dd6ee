// serve.hip -- the serving loop: a stream of independent solve requests of uneven sizes admitted
// by an HBM budget and solved with continuous batching (no reference counterpart: the reference
// solves one system per call).
//
// A request is an assembled square matrix, a right-hand side and an output vector (holding the
// initial guess), all on the device.  Requests wait in arrival order.  Each call of step():
//   1. queues one graph replay (16 CG iterations of every running system) of every live group,
//      each on its own stream, so the groups' work overlaps on the device;
//   2. admits, in arrival order, the waiting requests that fit the HBM budget and the slot limit
//      as one new group -- a batch of their systems (block-diagonal copy, its own inner CG engine,
//      its own stream) -- and starts it: requests that arrive while others run join as soon as
//      memory and slots free up, without waiting for the running groups to finish;
//   3. completes every group's replay: the systems that paused get their reliable FP64 update,
//      and each system that has converged (or run out of budget) is retired at once -- its
//      solution is copied to the request's output and reported -- while the others run on;
//   4. frees the groups whose systems have all retired, returning their bytes to the budget.
// A system's arithmetic depends only on its own matrix and vectors (fixed SpMV tiling per system,
// per-system reductions in a fixed order), so its result is bit-identical whatever group it is
// admitted with and whatever else is running.
#include "batch.hpp"

#include <deque>
#include <list>

using namespace femd_amd;

namespace {

struct Request {
    int64_t tag = 0;
    femd_csr A = nullptr;
    const double *b = nullptr;
    double *x = nullptr;
    int64_t bytes = 0;
};

struct Group {
    std::vector<Request> reqs;
    std::unique_ptr<femd_batch_s> B;
    DevBuf bcat, xcat;
    hipStream_t stream = nullptr;
    int64_t bytes = 0;
    int left = 0;
    ~Group()
    {
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
            B.reset();                          // its graph belongs to the stream
            (void)hipStreamDestroy(stream);
        }
    }
};

struct Done {
    int64_t tag;
    int iterations;
    double residual;
    int converged;
};

} // namespace

struct femd_server_s {
    int64_t budget = 0, used = 0;
    int max_systems = 0, running = 0;
    int precision = FEMD_FP32;
    double tol = 1e-10, inner_tol = 0.3;
    int max_outer = 200;
    std::deque<Request> queue;
    std::list<std::unique_ptr<Group>> groups;
    std::deque<Done> done;                      // retired, not yet handed out
};

namespace {

// Device bytes a request holds while admitted: the batch's copy of its matrix (pattern, values,
// tile plan, scaled values and their dictionary) and the solver's vectors.
int64_t request_bytes(const femd_csr_s *A, int precision)
{
    const int64_t n = A->pat->n, nnz = A->pat->nnz;
    const int64_t w = precision == FEMD_FP64 ? 8 : 4;
    const int64_t per_entry = 8 + 4 + 2 + 2 + 1 + 4 + (precision == FEMD_FP64 ? 8 : 0);
    const int64_t per_row = 4 + 5 * w + 16 + 12 + 16 + 16;
    return per_entry * nnz + per_row * n + 65536;
}

void admit(femd_server_s *S)
{
    std::vector<Request> take;
    int64_t bytes = 0;
    while (!S->queue.empty() && S->running + static_cast<int>(take.size()) < S->max_systems &&
           S->used + bytes + S->queue.front().bytes <= S->budget)
    {
        take.push_back(S->queue.front());
        bytes += S->queue.front().bytes;
        S->queue.pop_front();
    }
    if (take.empty()) return;
    auto G = std::make_unique<Group>();
    HIPC(hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking));
    G->reqs = take;
    G->bytes = bytes;
    G->left = static_cast<int>(take.size());
    std::vector<femd_csr> mats;
    for (const Request &r : take) mats.push_back(r.A);
    // the group's engine works in the inner CG's order when that order and the tile lists are
    // built on the device; built on the host for every group they sat on the admission path and
    // cost the stream more than the faster iterations gave back (200 requests: 15.3 against 16.0
    // solves/s, p50 0.89 against 0.50 s)
    G->B = batch_build(static_cast<int>(mats.size()), mats.data(), G->stream, inner_order_enabled() && device_setup_enabled());
    const int64_t nr = G->B->nrows;
    G->bcat = DevBuf(8 * size_t(std::max<int64_t>(nr, 1)));
    G->xcat = DevBuf(8 * size_t(std::max<int64_t>(nr, 1)));
    for (size_t p = 0; p < take.size(); ++p)
    {
        const int64_t r0 = G->B->roff[p], m = G->B->roff[p + 1] - r0;
        if (m == 0) continue;
        HIPC(hipMemcpyAsync(G->bcat.as<double>() + r0, take[p].b, 8 * size_t(m), hipMemcpyDeviceToDevice, G->stream));
        HIPC(hipMemcpyAsync(G->xcat.as<double>() + r0, take[p].x, 8 * size_t(m), hipMemcpyDeviceToDevice, G->stream));
    }
    CGEngine &E = G->B->engine[S->precision == FEMD_FP32 ? 0 : 1];
    E.setup(G->B->A.get(), S->precision, G->stream, false);
    E.begin(G->bcat.as<double>(), G->xcat.as<double>(), S->tol, S->inner_tol, S->max_outer, true, G->stream);
    S->used += bytes;
    S->running += static_cast<int>(take.size());
    S->groups.push_back(std::move(G));
}

// hand out system p of group G: its solution to the request's output (copied on the group's
// stream; the caller synchronizes that stream before the result can be handed out), its result to
// the list
void retire(femd_server_s *S, Group &G, int p)
{
    const CGEngine &E = G.B->engine[S->precision == FEMD_FP32 ? 0 : 1];
    const int64_t r0 = G.B->roff[size_t(p)], m = G.B->roff[size_t(p) + 1] - r0;
    if (m > 0)
        HIPC(hipMemcpyAsync(G.reqs[size_t(p)].x, G.xcat.as<double>() + r0, 8 * size_t(m), hipMemcpyDeviceToDevice, G.stream));
    Done d{G.reqs[size_t(p)].tag, 0, 0.0, 0};
    E.system_result(p, &d.iterations, &d.residual, &d.converged);
    S->done.push_back(d);
    G.left -= 1;
    S->running -= 1;
}

void step(femd_server_s *S)
{
    const int eng = S->precision == FEMD_FP32 ? 0 : 1;
    for (auto &G : S->groups) G->B->engine[eng].launch(G->stream);
    const size_t before = S->groups.size();
    admit(S);
    // a new group's systems that begin() already retired (zero right-hand side, or the guess
    // already converged)
    if (S->groups.size() > before)
    {
        Group &G = *S->groups.back();
        CGEngine &E = G.B->engine[eng];
        for (int p = 0; p < G.B->np; ++p)
        {
            int it = 0, conv = 0;
            double res = 0.0;
            E.system_result(p, &it, &res, &conv);
            if (conv && it == 0) retire(S, G, p);
        }
        E.launch(G.stream);
    }
    for (auto &G : S->groups)
    {
        CGEngine &E = G->B->engine[eng];
        const size_t ndone = S->done.size();
        for (int p : E.complete(G->stream)) retire(S, *G, p);
        // a retired request's output must hold its solution before step() reports it, while the
        // group's other systems run on: wait for the copies retire() queued
        if (S->done.size() != ndone) HIPC(hipStreamSynchronize(G->stream));
    }
    for (auto it = S->groups.begin(); it != S->groups.end();)
    {
        Group &G = **it;
        if (G.left == 0 && G.B->engine[eng].finished())
        {
            HIPC(hipStreamSynchronize(G.stream));        // the output copies
            S->used -= G.bytes;
            it = S->groups.erase(it);
        }
        else ++it;
    }
}

femd_server_s *check_server(femd_server S)
{
    if (!S) fail(FEMD_EINVAL, "null server");
    return S;
}

} // namespace

extern "C" {

int femd_server_create(int64_t hbm_budget, int max_systems, int precision, double tol, double inner_tol, int max_outer,
                       femd_server *out)
{
    return api([&] {
        if (!out) fail(FEMD_EINVAL, "server: null argument");
        if (hbm_budget <= 0 || max_systems < 1) fail(FEMD_EINVAL, "server: need hbm_budget > 0 and max_systems >= 1");
        if (precision != FEMD_FP32 && precision != FEMD_FP64) fail(FEMD_EINVAL, "server: precision is FEMD_FP32 or FEMD_FP64");
        if (!(tol >= 0.0) || !(inner_tol > 0.0 && inner_tol < 1.0) || max_outer < 1)
            fail(FEMD_EINVAL, "server: need tol >= 0, 0 < inner_tol < 1, max_outer >= 1");
        ensure_device();
        auto S = std::make_unique<femd_server_s>();
        S->budget = hbm_budget;
        S->max_systems = max_systems;
        S->precision = precision;
        S->tol = tol;
        S->inner_tol = inner_tol;
        S->max_outer = max_outer;
        *out = S.release();
    });
}

int femd_server_submit(femd_server S, int64_t tag, femd_csr A, const double *b_dev, double *x_dev, int64_t *bytes)
{
    return api([&] {
        femd_server_s *sv = check_server(S);
        femd_csr_s *a = check_csr(A);
        if (a->pat->n != a->pat->m) fail(FEMD_EINVAL, "server: the matrix must be square");
        if (a->pat->n > 0 && (!b_dev || !x_dev)) fail(FEMD_EINVAL, "server: null vector");
        Request r{tag, A, b_dev, x_dev, request_bytes(a, sv->precision)};
        if (r.bytes > sv->budget)
            fail(FEMD_EINVAL, "server: the request needs " + std::to_string(r.bytes) + " bytes, more than the budget of " +
                                  std::to_string(sv->budget));
        if (bytes) *bytes = r.bytes;
        sv->queue.push_back(r);
    });
}

int femd_server_step(femd_server S, int max_done, int64_t *tags, int *iterations, double *residuals, int *converged,
                     int *ndone)
{
    return api([&] {
        femd_server_s *sv = check_server(S);
        if (!ndone || max_done < 0 || (max_done > 0 && !tags)) fail(FEMD_EINVAL, "server: bad result buffers");
        if (sv->done.empty()) step(sv);
        int k = 0;
        while (k < max_done && !sv->done.empty())
        {
            const Done &d = sv->done.front();
            tags[k] = d.tag;
            if (iterations) iterations[k] = d.iterations;
            if (residuals) residuals[k] = d.residual;
            if (converged) converged[k] = d.converged;
            sv->done.pop_front();
            ++k;
        }
        *ndone = k;
    });
}

int femd_server_status(femd_server S, int *queued, int *running, int *groups, int64_t *bytes_in_use)
{
    return api([&] {
        femd_server_s *sv = check_server(S);
        if (queued) *queued = static_cast<int>(sv->queue.size());
        if (running) *running = sv->running;
        if (groups) *groups = static_cast<int>(sv->groups.size());
        if (bytes_in_use) *bytes_in_use = sv->used;
    });
}

void femd_server_free(femd_server S)
{
    delete S;
}

} // extern "C"
