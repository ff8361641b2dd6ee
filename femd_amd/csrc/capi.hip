// capi.hip -- the C ABI of include/femd_hip.h: handles, the device memory pool, and the
// orchestration of the kernels in kernels.hip.
#include "cg_engine.hpp"
#include "sweep.hpp"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <thread>

using namespace femd_amd;

namespace femd_amd {

thread_local std::string g_err;
Device g_dev;
std::mutex g_init_mu;

// ---------------------------------------------------------------------- no-allocation work space
// The device work space of femd_cg_dev_noalloc: the CG's four vectors of at most NOALLOC_MAX_ROWS
// entries, its two arrays of chunk partials, eight scalars and the loop state (32 MiB and a few KiB
// per GPU, held from the library's first use on that GPU until the process ends).  In host code
// `scratch` only names the array: its device address on each GPU comes from hipGetSymbolAddress.
constexpr int64_t NOALLOC_MAX_ROWS = int64_t(1) << 20;
constexpr int64_t NOALLOC_MAX_CHUNKS = (NOALLOC_MAX_ROWS + CHUNK - 1) / CHUNK;
constexpr int64_t NOALLOC_STATE = (sizeof(CGState) + 7) / 8 + 8;
constexpr int64_t NOALLOC_DOUBLES = 4 * NOALLOC_MAX_ROWS + 2 * NOALLOC_MAX_CHUNKS + 8 + NOALLOC_STATE;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};              // scratch's device address on each GPU
std::string g_scratch_err[MAX_DEVICES];           // why it is missing, when hipGetSymbolAddress failed

// Called once per device, with that device current (from the library's set-up).
void noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES || g_scratch[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
    {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_scratch[device] = static_cast<double *>(ptr);
}

void ensure_device()
{
    if (g_dev.ready) return;
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_dev.ready) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) fail(FEMD_EDEVICE, "no HIP device is visible");
    HIPC(hipSetDevice(g_dev.id));
    HIPC(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
    noalloc_setup(g_dev.id);
    g_dev.ready = true;
}

void *pool_alloc(size_t bytes)
{
    ensure_device();
    const size_t sz = round_size(bytes);
    {
        std::lock_guard<std::mutex> lk(g_dev.mu);
        auto it = g_dev.free_blocks.find(sz);
        if (it != g_dev.free_blocks.end())
        {
            void *p = it->second;
            g_dev.free_blocks.erase(it);
            g_dev.cached -= sz;
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess)
    {
        {   // release the cache once and retry
            std::lock_guard<std::mutex> lk(g_dev.mu);
            (void)hipDeviceSynchronize();
            for (auto &kv : g_dev.free_blocks) (void)hipFree(kv.second);
            g_dev.free_blocks.clear();
            g_dev.cached = 0;
        }
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) fail(FEMD_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return p;
}

void pool_free(void *p, size_t bytes)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_dev.mu);
    g_dev.free_blocks.emplace(round_size(bytes), p);
    g_dev.cached += round_size(bytes);
}

std::shared_ptr<DevPattern> make_pattern(int n, int m, std::vector<int> ip, std::vector<int> ix, hipStream_t s)
{
    auto P = std::make_shared<DevPattern>();
    P->n = n; P->m = m; P->nnz = ip.empty() ? 0 : ip.back();
    P->h_ip = std::move(ip); P->h_ix = std::move(ix);
    P->ip = upload(P->h_ip, s);
    P->ix = upload(P->h_ix, s);
    return P;
}

const int *diag_of(DevPattern &P, hipStream_t s)
{
    if (!P.have_diag)
    {
        P.diag = DevBuf(sizeof(int) * std::max(P.n, 1));
        launch_diag_positions(P.n, P.ip.as<int>(), P.ix.as<int>(), P.diag.as<int>(), s);
        P.have_diag = true;
    }
    return P.diag.as<int>();
}

const float *values32(femd_csr_s *A, hipStream_t s)
{
    if (!A->v32_valid)
    {
        A->v32.ensure(sizeof(float) * std::max<int64_t>(A->pat->nnz, 1));
        launch_to_float(A->pat->nnz, A->v.as<double>(), A->v32.as<float>(), s);
        A->v32_valid = true;
    }
    return A->v32.as<float>();
}

Quad &quad_of(femd_space_s *V, int degree, hipStream_t s)
{
    auto it = V->quads.find(degree);
    if (it != V->quads.end()) return *it->second;
    auto Q = std::make_unique<Quad>();
    const bool quad = V->V.nv == 4;
    QuadRule R = quad ? square_rule(degree) : triangle_rule(degree);
    Q->nq = R.size();
    Q->xi = R.xi; Q->eta = R.eta; Q->w = R.w;
    const int n = V->V.nloc;
    std::vector<double> tab(static_cast<size_t>(Q->nq) * 3 * n);
    for (int q = 0; q < Q->nq; ++q)
    {
        double *o = &tab[static_cast<size_t>(q) * 3 * n];
        if (quad) V->refq->tabulate(R.xi[q], R.eta[q], o);
        else V->ref->tabulate(R.xi[q], R.eta[q], o);
    }
    Q->tab = upload(tab, s);
    Q->qw = upload(R.w, s);
    Q->qxi = upload(R.xi, s);
    Q->qeta = upload(R.eta, s);
    HIPC(hipStreamSynchronize(s));
    auto &ref = *Q;
    V->quads[degree] = std::move(Q);
    return ref;
}

void ensure_host_pattern(femd_space_s *V)
{
    if (!V->have_host_pattern) { V->P = build_pattern(V->V); V->have_host_pattern = true; }
}

void ensure_geom(femd_space_s *V, hipStream_t s)
{
    if (!V->geom.p) { V->geom = upload(V->V.geom, s); HIPC(hipStreamSynchronize(s)); }
}

void ensure_pattern(femd_space_s *V, hipStream_t s)
{
    if (V->have_pattern) return;
    if (!V->have_host_pattern && device_setup_enabled())
    {
        // built on the device from cell_dof; the host keeps the row pointers and columns
        ensure_geom(V, s);
        if (!V->cell_dof_dev.p) V->cell_dof_dev = upload(V->V.cell_dof, s);
        DevPatternLists D;
        device_build_pattern(V->V.dim, V->V.ncells, V->V.nloc, V->cell_dof_dev.as<int>(), D, s);
        V->P = PatternData{};
        V->P.n = V->V.dim;
        V->P.indptr = D.h_ip;
        V->P.indices = D.h_ix;
        V->have_host_pattern = true;
        auto P = std::make_shared<DevPattern>();
        P->n = P->m = V->V.dim;
        P->nnz = D.nnz;
        P->h_ip = std::move(D.h_ip);
        P->h_ix = std::move(D.h_ix);
        P->ip = std::move(D.ip);
        P->ix = std::move(D.ix);
        V->pat = std::move(P);
        V->mptr = std::move(D.mptr);
        V->mlist = std::move(D.mlist);
        V->vptr = std::move(D.vptr);
        V->vlist = std::move(D.vlist);
        V->have_pattern = true;
        return;
    }
    ensure_host_pattern(V);
    ensure_geom(V, s);
    V->pat = make_pattern(V->V.dim, V->V.dim, V->P.indptr, V->P.indices, s);
    V->mptr = upload(V->P.mptr, s);
    V->mlist = upload(V->P.mlist, s);
    V->vptr = upload(V->P.vptr, s);
    V->vlist = upload(V->P.vlist, s);
    HIPC(hipStreamSynchronize(s));
    V->have_pattern = true;
}

// The coefficient arrays of the terms: device arrays [cell * nq + q], or constants.
struct Coeffs {
    std::vector<DevBuf> bufs;
    TermSet T{};
};

void check_program(const femd_coefficient &c, bool normals = false)
{
    if (c.ninstr < 1) fail(FEMD_EINVAL, "coefficient: empty program");
    if (c.nreg < 1) fail(FEMD_EINVAL, "coefficient: at least one register");
    for (int k = 0; k < c.ninstr; ++k)
    {
        const int op = c.code[4 * k], o = c.code[4 * k + 1], a = c.code[4 * k + 2], b = c.code[4 * k + 3];
        if (o < 0 || o >= c.nreg) fail(FEMD_EINVAL, "coefficient: register out of range");
        const bool unary = op >= 9 && op <= 19;
        if (((op == 3 || op == 4) && !normals) || (op == 5 && normals) || op < 0 || op > 19)
            fail(FEMD_EINVAL, "coefficient: instruction " + std::to_string(op) +
                                  (op == 5 ? " (a known field) is supported in dx terms only"
                                           : " (a facet normal) needs a ds term"));
        if (op == 5 && (a < 0 || a >= c.ninputs || !c.inputs))
            fail(FEMD_EINVAL, "coefficient: input " + std::to_string(a) + " out of range (" + std::to_string(c.ninputs) +
                                  " known fields given)");
        if (op == 0 && (a < 0 || a >= c.nconst)) fail(FEMD_EINVAL, "coefficient: constant out of range");
        if ((op == 6 || op == 7) && (a < 0 || a >= c.nreg || b < 0 || b >= c.nreg)) fail(FEMD_EINVAL, "coefficient: register out of range");
        if (op == 8 && (a < 0 || a >= c.nreg || b < 0 || b >= c.nconst)) fail(FEMD_EINVAL, "coefficient: pow operands out of range");
        if (unary && (a < 0 || a >= c.nreg)) fail(FEMD_EINVAL, "coefficient: register out of range");
    }
}

// Register renaming by liveness: every instruction's result gets a register that is free from
// its definition to its last use, so a program of any size runs in a few registers.  The
// operations and their order are unchanged.
std::vector<int> compact_registers(const femd_coefficient &c, int &nreg_out)
{
    const int n = c.ninstr;
    std::vector<int> code(c.code, c.code + 4 * static_cast<size_t>(n));
    auto reads = [&](int k, int slot) {            // slot 0: a, 1: b; is that operand a register?
        const int op = code[4 * k];
        if (slot == 0) return op >= 6 && op <= 19;
        return op == 6 || op == 7;
    };
    std::vector<int> writer(static_cast<size_t>(c.nreg), -1);     // last instruction that wrote a register
    std::vector<int> src(2 * static_cast<size_t>(n), -1);          // the instruction whose value an operand reads
    std::vector<int> last(static_cast<size_t>(n), -1);             // last reader of each instruction's value
    for (int k = 0; k < n; ++k)
    {
        for (int slot = 0; slot < 2; ++slot)
            if (reads(k, slot))
            {
                const int w = writer[code[4 * k + 2 + slot]];
                if (w < 0) fail(FEMD_EINVAL, "coefficient: a register is read before it is written");
                src[2 * k + slot] = w;
                last[w] = k;
            }
        writer[code[4 * k + 1]] = k;
    }
    last[n - 1] = n;                                               // the result lives to the end
    std::vector<int> reg_of(static_cast<size_t>(n), -1), free_regs;
    int nreg = 0;
    for (int k = 0; k < n; ++k)
    {
        for (int slot = 0; slot < 2; ++slot)
            if (src[2 * k + slot] >= 0) code[4 * k + 2 + slot] = reg_of[src[2 * k + slot]];
        for (int slot = 0; slot < 2; ++slot)
        {
            const int w = src[2 * k + slot];
            if (w >= 0 && last[w] == k && !(slot == 1 && src[2 * k] == w)) free_regs.push_back(reg_of[w]);
        }
        int r;
        if (!free_regs.empty()) { r = free_regs.back(); free_regs.pop_back(); }
        else r = nreg++;
        reg_of[k] = r;
        code[4 * k + 1] = r;
        if (last[k] < 0) free_regs.push_back(r);                   // a value nobody reads
    }
    nreg_out = nreg;
    return code;
}

// The known fields a program reads, at the rule's points of V's cells (one device array each).
std::vector<DevBuf> eval_inputs(femd_space_s *V, int degree, const femd_coefficient &c, InputSet &in, hipStream_t s)
{
    if (c.ninputs > MAX_INPUTS) fail(FEMD_EINVAL, "coefficient: at most " + std::to_string(MAX_INPUTS) + " known fields");
    std::vector<DevBuf> out;
    in = InputSet{};
    for (int i = 0; i < c.ninputs; ++i)
    {
        const femd_field &f = c.inputs[i];
        femd_space_s *W = f.space;
        if (!W || !f.coefficients) fail(FEMD_EINVAL, "known field: null space or coefficients");
        if (W->V.ncells != V->V.ncells || W->V.nv != V->V.nv)
            fail(FEMD_EINVAL, "known field: its space is not on the form's cells");
        if (f.code < 0 || f.code > 2) fail(FEMD_EINVAL, "derivative code is 0 (value), 1 (d/dx) or 2 (d/dy)");
        ensure_geom(W, s);
        if (!W->cell_dof_dev.p) W->cell_dof_dev = upload(W->V.cell_dof, s);
        Quad &QW = quad_of(W, degree, s);
        DevBuf u = upload(f.coefficients, size_t(std::max(W->V.dim, 1)), s);
        out.emplace_back(8 * std::max<size_t>(size_t(V->V.ncells) * QW.nq, 1));
        if (!launch_eval_field(W->V.nv, W->V.nloc, W->V.ncells, QW.nq, W->cell_dof_dev.as<int>(), u.as<double>(),
                               QW.tab.as<double>(), W->geom.as<double>(), QW.qxi.as<double>(), QW.qeta.as<double>(), f.code,
                               out.back().as<double>(), s))
            fail(FEMD_EINVAL, "known field: no kernel for its local size");
        in.p[i] = out.back().as<double>();
        HIPC(hipStreamSynchronize(s));       // u is released on return
    }
    return out;
}

void eval_program(femd_space_s *V, Quad &Q, int degree, const femd_coefficient &c, double *out, hipStream_t s)
{
    check_program(c);
    int nreg = 0;
    std::vector<int> compact = compact_registers(c, nreg);
    if (nreg > MAX_REGS) fail(FEMD_EINVAL, "coefficient: the program needs more than " + std::to_string(MAX_REGS) + " live registers");
    InputSet in;
    std::vector<DevBuf> fields = eval_inputs(V, degree, c, in, s);
    DevBuf code = upload(compact, s);
    DevBuf consts = upload(c.consts, static_cast<size_t>(std::max(c.nconst, 0)), s);
    launch_coefficient(V->V.ncells, Q.nq, V->V.nv, V->geom.as<double>(), Q.qxi.as<double>(), Q.qeta.as<double>(), c.ninstr,
                       code.as<int>(), nreg, consts.as<double>(), in, out, s);
    HIPC(hipStreamSynchronize(s));           // code, consts and the field values are released on return
}

// Two coefficients with the same host values array, or the same program (instructions and
// constants), have the same values at every point.
bool same_coefficient(const femd_coefficient &x, const femd_coefficient &y)
{
    if (x.values || y.values) return x.values == y.values;
    if (x.ninstr != y.ninstr || x.nconst != y.nconst || x.ninputs != y.ninputs) return false;
    for (int i = 0; i < x.ninputs; ++i)
        if (x.inputs[i].space != y.inputs[i].space || x.inputs[i].coefficients != y.inputs[i].coefficients ||
            x.inputs[i].code != y.inputs[i].code)
            return false;
    return std::equal(x.code, x.code + 4 * x.ninstr, y.code) && std::equal(x.consts, x.consts + x.nconst, y.consts);
}

// The terms' coefficients on the device: a constant stays a value; a program is evaluated at the
// rule's points (once for terms whose coefficients are the same, which then share the array).
Coeffs make_coeffs(femd_space_s *V, Quad &Q, int degree, int nt, const int *a, const int *b, const femd_coefficient *c,
                   hipStream_t s)
{
    if (nt < 1 || nt > MAX_TERMS) fail(FEMD_EINVAL, "assembly: 1.." + std::to_string(MAX_TERMS) + " terms per call");
    // the element kernels stage the basis table and the weights in 64 KiB of LDS
    if (sizeof(double) * (size_t(Q.nq) * 3 * V->V.nloc + Q.nq) > 65536)
        fail(FEMD_EINVAL, "assembly: the quadrature rule of this degree is too large for the element kernels");
    Coeffs C;
    C.T.nt = nt;
    const size_t np = static_cast<size_t>(V->V.ncells) * Q.nq;
    for (int t = 0; t < nt; ++t)
    {
        if (a[t] < 0 || a[t] > 2 || (b && (b[t] < 0 || b[t] > 2)))
            fail(FEMD_EINVAL, "assembly_2d: derivative code is not in 0..2");
        C.T.a[t] = a[t];
        C.T.b[t] = b ? b[t] : 0;
        C.T.c[t] = nullptr;
        C.T.cval[t] = 0.0;
        const femd_coefficient &k = c[t];
        int twin = -1;
        for (int u = 0; u < t && twin < 0; ++u)
            if (C.T.c[u] && same_coefficient(k, c[u])) twin = u;
        if (twin >= 0)
        {
            C.T.c[t] = C.T.c[twin];
            continue;
        }
        if (k.values)
        {
            C.bufs.push_back(upload(k.values, np, s));
            C.T.c[t] = C.bufs.back().as<double>();
        }
        else if (k.ninstr == 1 && k.code[0] == 0)
        {
            if (k.code[2] < 0 || k.code[2] >= k.nconst) fail(FEMD_EINVAL, "coefficient: constant out of range");
            C.T.cval[t] = k.consts[k.code[2]];
        }
        else
        {
            C.bufs.emplace_back(np * sizeof(double));
            eval_program(V, Q, degree, k, C.bufs.back().as<double>(), s);
            C.T.c[t] = C.bufs.back().as<double>();
        }
    }
    return C;
}

femd_space_s *check_space(femd_space V)
{
    if (!V) fail(FEMD_EINVAL, "null space");
    return V;
}
// y = A x in FP64 with the reference's row order (the tuned CSR-stream variant).
void spmv64(femd_csr_s *A, const double *x, double *y, const int *stop, hipStream_t s)
{
    DevPattern &P = *A->pat;
    spmv_run64(spmv_plan(P, FEMD_FP64, A->v.p, s), P, A->v.as<double>(), x, y, stop, s);
}

// The ordered dot product into a device scalar.
void ddot_dev(int64_t n, const double *x, const double *y, double *parts, double *out, hipStream_t s)
{
    launch_dot_parts(n, x, y, parts, nullptr, s);
    launch_sum_parts(nchunks(n), parts, out, s);
}

// z = M^{-1} r on the device (any kind), stop may be null.
void precond_apply(femd_precond_s *M, const double *r, double *z, const int *stop, hipStream_t s)
{
    const int n = M->n;
    if (M->kind == FEMD_PC_JACOBI) { launch_mul(n, M->inv.as<double>(), r, z, s); return; }
    const bool perm = M->ordering == FEMD_ORDER_RCM;
    if (M->sweep)
    {
        // the planned sweeps: r (permuted) into the forward stream, the forward sweep writes y into
        // the backward stream, the backward sweep writes z (scattered back through the order)
        const int ilu = M->kind == FEMD_PC_ILU0;
        double *stage = M->swork.as<double>();
        launch_sweep_fill(n, M->inposF.as<int>(), perm ? M->perm.as<int>() : nullptr, r, M->streamF.as<double>(), stop, s);
        launch_sweep(ilu ? SW_ILU_FWD : SW_SSOR_FWD, M->eregF, M->nstepsF, M->streamF.p, M->firstF.p, M->zunitF, M->omega,
                     stage, stop, s);
        launch_sweep_scatter(n, M->scatF.as<int>(), stage, M->streamB.as<double>(), stop, s);
        launch_sweep(ilu ? SW_ILU_BWD : SW_SSOR_BWD, M->eregB, M->nstepsB, M->streamB.p, M->firstB.p, M->zunitB, M->omega,
                     stage, stop, s);
        launch_sweep_scatter(n, M->scatB.as<int>(), stage, z, stop, s);
        return;
    }
    const double *rin = r;
    double *zout = z;
    if (perm)
    {
        launch_gather_perm(n, M->perm.as<int>(), r, M->rp.as<double>(), s);
        rin = M->rp.as<double>();
        zout = M->zp.as<double>();
    }
    if (M->kind == FEMD_PC_ILU0)
        launch_ilu0_apply(n, M->nlevL, M->lptrL.as<int>(), M->lrowsL.as<int>(), M->nlevU, M->lptrU.as<int>(), M->lrowsU.as<int>(),
                          M->ip.as<int>(), M->ix.as<int>(), M->diag.as<int>(), M->lu.as<double>(), rin, zout, stop, s);
    else
        launch_ssor_apply(n, M->nlevL, M->lptrL.as<int>(), M->lrowsL.as<int>(), M->nlevU, M->lptrU.as<int>(), M->lrowsU.as<int>(),
                          M->ip.as<int>(), M->ix.as<int>(), M->diag.as<int>(), M->lu.as<double>(), M->omega, rin, zout, stop, s);
    if (perm) launch_scatter_perm(n, M->perm.as<int>(), M->zp.as<double>(), z, s);
}

// ------------------------------------------------------------------------------------ CG
// krylov/cg.hpp:245-350.  The set-up and the final residual are driven from the host (a few
// read-backs per solve); the iterations run from device-side scalars (CGState), a batch of
// iterations enqueued between read-backs, every kernel a no-op once `stop` is set.
// The device work space of one CG solve: four vectors of n (z only with a general preconditioner),
// two arrays of chunk partials, eight scalars and the loop state.
struct CGWork {
    double *r, *p, *Ap, *z, *prr, *prz, *scal;
    CGState *st;
};

void cg_check(femd_csr_s *A, femd_precond_s *M)
{
    const int n = A->pat->n;
    if (A->pat->m != n) fail(FEMD_EINVAL, "cg: the matrix must be square");
    if (M && M->n != n) fail(FEMD_EINVAL, "krylov: preconditioner and right-hand side differ in size");
}

// The solve on a given work space; hst receives the loop state at every read-back (pinned or not).
void cg_run(femd_csr_s *A, femd_precond_s *M, const double *b, double *x, int max_iter, double tol,
            femd_solve_info *info, const CGWork &w, CGState *hst, hipStream_t s)
{
    const int n = A->pat->n;
    std::memset(info, 0, sizeof(*info));
    const bool jac = M && M->kind == FEMD_PC_JACOBI, general = M && !jac;
    const int np = nchunks(n);
    double *rd = w.r, *pd = w.p, *Apd = w.Ap, *zd = w.z, *prr = w.prr, *prz = w.prz, *sc = w.scal;
    CGState *st = w.st;
    const double *inv = jac ? M->inv.as<double>() : nullptr;

    ddot_dev(n, b, b, prr, sc, s);
    const double nb = std::sqrt(read_scalar(sc, s));
    if (nb == 0.0) { launch_fill(n, 0.0, x, s); HIPC(hipStreamSynchronize(s)); info->converged = 1; return; }
    spmv64(A, x, Apd, nullptr, s);
    launch_sub(n, b, Apd, rd, s);
    ddot_dev(n, rd, rd, prr, sc, s);
    double rn = std::sqrt(read_scalar(sc, s));
    if (rn <= tol * nb) { info->converged = 1; info->residual = rn / nb; info->precond_residual = info->residual; return; }
    double rz;
    if (general)
    {
        precond_apply(M, rd, zd, nullptr, s);
        launch_copy(n, zd, pd, s);
        ddot_dev(n, rd, zd, prr, sc, s);
        rz = read_scalar(sc, s);
    }
    else if (jac)
    {
        launch_mul(n, inv, rd, pd, s);
        ddot_dev(n, rd, pd, prr, sc, s);
        rz = read_scalar(sc, s);
    }
    else { launch_copy(n, rd, pd, s); rz = rn * rn; }

    CGState h{};
    h.nb = nb; h.tolnb = tol * nb; h.rz = rz; h.it = 0; h.max_iter = max_iter;
    if (max_iter <= 0) h.stop = 1;
    else if (!(rz > 0.0)) { h.stop = 1; h.breakdown = 1; }
    HIPC(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    const int *stop = &st->stop;
    int batch = 8;
    while (!h.stop)
    {
        for (int k = 0; k < batch; ++k)
        {
            spmv64(A, pd, Apd, stop, s);
            launch_dot_parts(n, pd, Apd, prr, stop, s);
            launch_cg_alpha(prr, np, st, s);
            launch_cg_update(n, x, rd, pd, Apd, inv, prr, prz, st, s);
            launch_cg_beta(prr, prz, np, general ? 2 : (jac ? 1 : 0), st, s);
            if (general)
            {
                precond_apply(M, rd, zd, stop, s);
                launch_dot_parts(n, rd, zd, prr, stop, s);
                launch_cg_rz(prr, np, st, s);
                launch_cg_pupdate(n, pd, zd, nullptr, st, s);
            }
            else launch_cg_pupdate(n, pd, rd, inv, st, s);
        }
        HIPC(hipMemcpyAsync(hst, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        std::memcpy(&h, hst, sizeof(h));
        batch = std::min(64, batch * 2);
    }
    // the true residual and the verdict (cg.hpp:222-237)
    spmv64(A, x, Apd, nullptr, s);
    launch_sub(n, b, Apd, rd, s);
    ddot_dev(n, rd, rd, prr, sc, s);
    const double res = std::sqrt(read_scalar(sc, s)) / nb;
    info->iterations = h.it;
    info->breakdown = h.breakdown;
    info->residual = res;
    info->precond_residual = res;
    info->converged = !h.breakdown && res <= tol * (1.0 + 1e-6) + 1e-15;
    if (!info->converged && !h.breakdown && res <= 10.0 * tol && h.it < max_iter) info->converged = 1;
}

// The solve on buffers of the library's pool.
void cg_device(femd_csr_s *A, femd_precond_s *M, const double *b, double *x, int max_iter, double tol,
               femd_solve_info *info, hipStream_t s)
{
    cg_check(A, M);
    const int n = A->pat->n, np = nchunks(n);
    const bool general = M && M->kind != FEMD_PC_JACOBI;
    DevBuf r(8 * size_t(n)), p(8 * size_t(n)), Ap(8 * size_t(n)), z(general ? 8 * size_t(n) : 0);
    DevBuf prr(8 * size_t(std::max(np, 1))), prz(8 * size_t(std::max(np, 1))), scal(64), stb(sizeof(CGState));
    Pinned hst(sizeof(CGState));
    const CGWork w{r.as<double>(), p.as<double>(), Ap.as<double>(), z.as<double>(), prr.as<double>(), prz.as<double>(),
                   scal.as<double>(), stb.as<CGState>()};
    cg_run(A, M, b, x, max_iter, tol, info, w, hst.as<CGState>(), s);
}

// The solve on the device's `scratch` array: no allocation of any kind (the loop state is read
// back into host stack memory).
void cg_noalloc(femd_csr_s *A, femd_precond_s *M, const double *b, double *x, int max_iter, double tol,
                femd_solve_info *info, hipStream_t s)
{
    cg_check(A, M);
    const int64_t n = A->pat->n;
    if (n > NOALLOC_MAX_ROWS)
        fail(FEMD_EINVAL, "cg_dev_noalloc: " + std::to_string(n) + " rows, more than its work space holds (" +
                              std::to_string(NOALLOC_MAX_ROWS) + ")");
    if (n > 0 && !A->pat->plan64)
        fail(FEMD_EINVAL, "cg_dev_noalloc: the matrix has no FP64 SpMV launch shape yet; call "
                          "femd_spmv_tuning(A, FEMD_FP64, ...) (or any FP64 product or solve on A) once first");
    double *base = g_scratch[g_dev.id];
    if (!base) fail(FEMD_EDEVICE, "cg_dev_noalloc: no work space on this device: " + g_scratch_err[g_dev.id]);
    double *parts = base + 4 * NOALLOC_MAX_ROWS;
    const CGWork w{base, base + NOALLOC_MAX_ROWS, base + 2 * NOALLOC_MAX_ROWS, base + 3 * NOALLOC_MAX_ROWS, parts,
                   parts + NOALLOC_MAX_CHUNKS, parts + 2 * NOALLOC_MAX_CHUNKS,
                   reinterpret_cast<CGState *>(parts + 2 * NOALLOC_MAX_CHUNKS + 8)};
    CGState hst{};
    cg_run(A, M, b, x, max_iter, tol, info, w, &hst, s);
}

// --------------------------------------------------------------------------- GMRES / LGMRES
// krylov.hip: the solvers with their Arnoldi steps on the device
void krylov_device(femd_csr_s *A, femd_precond_s *M, const double *b, double *x, int lgmres, int restart, int max_iter,
                   double tol, int k_aug, femd_solve_info *info, hipStream_t s);

// ------------------------------------------------------------------------ mixed precision
void refined_device(femd_csr_s *A, const double *b, double *x, double tol, double inner_tol, int max_outer,
                    femd_solve_info *info, hipStream_t s)
{
    std::memset(info, 0, sizeof(*info));
    if (!(inner_tol > 0.0 && inner_tol < 1.0)) fail(FEMD_EINVAL, "refined solve: need 0 < inner_tol < 1");
    if (A->pat->n == 0) { info->converged = 1; return; }
    CGEngine cg;
    cg.setup(A, FEMD_FP32, s);
    double res = 0.0;
    const CGEngine::Result r = cg.solve(b, x, tol, inner_tol, std::max(max_outer, 1), true, nullptr, &res, s);
    info->converged = r.converged;
    info->iterations = r.iterations;
    info->breakdown = r.updates;
    info->residual = res;
    info->precond_residual = res;
}

// Plan the one-workgroup sweeps of an ILU(0) / SSOR preconditioner on its (permuted) pattern and
// pack its values (M->lu: the factor, or A's values for SSOR) into the streams.  When the schedule
// does not fit the LDS the level-set kernels stay in use (M->sweep false, the reason kept).
void plan_sweeps(femd_precond_s *M, const std::vector<int> &pip, const std::vector<int> &pix, const std::vector<int> &perm,
                 hipStream_t s)
{
    const int n = M->n;
    if (n == 0 || std::getenv("FEMD_LEVEL_SWEEPS")) return;
    std::vector<int> dg(n, -1);
    for (int i = 0; i < n; ++i)
        for (int p = pip[i]; p < pip[i + 1]; ++p)
            if (pix[p] == i) dg[i] = p;
    const bool ilu = M->kind == FEMD_PC_ILU0;
    SweepPlan B, F;
    std::string why, whyF;
    // the two sweeps' plans are independent: the backward one on a thread of its own
    bool okB = false;
    std::thread tb([&] { okB = plan_sweep(ilu ? SW_ILU_BWD : SW_SSOR_BWD, n, pip.data(), pix.data(), dg.data(), SW_ROWS, B, why); });
    const bool okF = plan_sweep(ilu ? SW_ILU_FWD : SW_SSOR_FWD, n, pip.data(), pix.data(), dg.data(), SW_ROWS, F, whyF);
    tb.join();
    bool ok = okB && okF;
    if (okB && !okF) why = whyF;
    if (ok && (F.stream.size() >= (size_t(1) << 31) || B.stream.size() >= (size_t(1) << 31)))
        ok = false, why = "a stream of 2 GiB or more";
    if (!ok) { M->sweep_note = why; return; }
    // where the sweeps' results go: the forward sweep's into the backward stream's vin fields, the
    // backward sweep's into the output vector (through the ordering)
    std::vector<int> scatF(n), scatB(n);
    for (int k = 0; k < n; ++k)
    {
        scatF[k] = int(B.in_pos[F.order[k]]);
        scatB[k] = perm.empty() ? B.order[k] : perm[B.order[k]];
    }
    const int64_t nnz = pip[n];
    std::vector<int> fmap(nnz), bmap(nnz), inpos(n);
    for (int64_t p = 0; p < nnz; ++p)
    {
        fmap[p] = int(F.vmap[p]);
        bmap[p] = int(B.vmap[p]);
    }
    for (int i = 0; i < n; ++i) inpos[i] = int(F.in_pos[i]);
    M->streamF = upload(F.stream, s);
    M->streamB = upload(B.stream, s);
    M->firstF = upload(F.first, s);
    M->firstB = upload(B.first, s);
    M->inposF = upload(inpos, s);
    M->scatF = upload(scatF, s);
    M->scatB = upload(scatB, s);
    M->swork = DevBuf(8 * size_t(n));
    {
        DevBuf df = upload(fmap, s), db = upload(bmap, s);
        launch_sweep_pack(nnz, df.as<int>(), db.as<int>(), M->lu.as<double>(), M->streamF.as<double>(), M->streamB.as<double>(), s);
        HIPC(hipStreamSynchronize(s));              // the maps go out of scope
    }
    M->nstepsF = F.nsteps;
    M->nstepsB = B.nsteps;
    M->zunitF = F.zunit;
    M->zunitB = B.zunit;
    M->eregF = F.ereg;
    M->eregB = B.ereg;
    M->nlevF = F.nlevels;
    M->nlevB = B.nlevels;
    M->bytesF = int64_t(F.stream.size());
    M->bytesB = int64_t(B.stream.size());
    M->entriesLU = F.entries + B.entries + n;
    M->sweep = true;
}

// Build the preconditioner's device data from A.
void build_precond(femd_precond_s *M, femd_csr_s *A, hipStream_t s)
{
    DevPattern &P = *A->pat;
    const int n = P.n;
    M->n = n;
    DevBuf bad(sizeof(int));
    const int big = 0x7fffffff;
    auto reset_bad = [&]() { HIPC(hipMemcpyAsync(bad.p, &big, sizeof(int), hipMemcpyHostToDevice, s)); };
    auto read_bad = [&]() {
        int h = 0;
        HIPC(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        return h;
    };
    if (M->kind == FEMD_PC_JACOBI)
    {
        M->inv = DevBuf(8 * size_t(std::max(n, 1)));
        reset_bad();
        launch_inv_diag(n, diag_of(P, s), A->v.as<double>(), M->inv.as<double>(), bad.as<int>(), s);
        const int b = read_bad();
        if (b != big) fail(FEMD_EZEROPIVOT, "Jacobi: zero diagonal entry in row " + std::to_string(b));
        return;
    }
    std::vector<int> perm;
    const int *hip_ = P.h_ip.data(), *hix = P.h_ix.data();
    std::vector<int> pip, pix;
    DevBuf src;
    if (M->ordering == FEMD_ORDER_RCM)
    {
        perm = rcm_order(n, hip_, hix);
        Permuted B = permute_pattern(n, hip_, hix, perm);
        pip = std::move(B.indptr); pix = std::move(B.indices);
        src = upload(B.src, s);
        M->perm = upload(perm, s);
        M->rp = DevBuf(8 * size_t(std::max(n, 1)));
        M->zp = DevBuf(8 * size_t(std::max(n, 1)));
    }
    else { pip = P.h_ip; pix = P.h_ix; }
    M->ip = upload(pip, s);
    M->ix = upload(pix, s);
    M->lu = DevBuf(8 * size_t(std::max<int64_t>(P.nnz, 1)));
    if (src.p) launch_take(P.nnz, src.as<int>(), A->v.as<double>(), M->lu.as<double>(), s);
    else if (P.nnz) HIPC(hipMemcpyAsync(M->lu.p, A->v.p, 8 * size_t(P.nnz), hipMemcpyDeviceToDevice, s));
    M->diag = DevBuf(4 * size_t(std::max(n, 1)));
    launch_diag_positions(n, M->ip.as<int>(), M->ix.as<int>(), M->diag.as<int>(), s);
    reset_bad();
    launch_inv_diag(n, M->diag.as<int>(), M->lu.as<double>(), nullptr, bad.as<int>(), s);
    const int b0 = read_bad();
    const char *what = M->kind == FEMD_PC_ILU0 ? "ILU(0)" : "SSOR";
    if (b0 != big) fail(FEMD_EZEROPIVOT, std::string(what) + ": zero diagonal entry in row " + std::to_string(b0));
    Levels L = lower_levels(n, pip.data(), pix.data()), U = upper_levels(n, pip.data(), pix.data());
    M->nlevL = L.count(); M->nlevU = U.count();
    M->lptrL = upload(L.ptr, s); M->lrowsL = upload(L.rows, s);
    M->lptrU = upload(U.ptr, s); M->lrowsU = upload(U.rows, s);
    if (M->kind == FEMD_PC_ILU0)
    {
        reset_bad();
        std::vector<int> hdg(n, -1);
        for (int i = 0; i < n; ++i)
            for (int p = pip[i]; p < pip[i + 1]; ++p)
                if (pix[p] == i) hdg[i] = p;
        if (!std::getenv("FEMD_ILU_ONE_GROUP") && ilu0_levels_fit(n, pip.data(), pix.data(), hdg.data()))
            launch_ilu0_levels(M->nlevL, L.ptr.data(), M->lrowsL.as<int>(), M->ip.as<int>(), M->ix.as<int>(),
                               M->diag.as<int>(), M->lu.as<double>(), bad.as<int>(), s);
        else
            launch_ilu0_factor(M->nlevL, M->lptrL.as<int>(), M->lrowsL.as<int>(), M->ip.as<int>(), M->ix.as<int>(),
                               M->diag.as<int>(), M->lu.as<double>(), bad.as<int>(), s);
        const int b1 = read_bad();
        if (b1 != big) fail(FEMD_EZEROPIVOT, "ILU(0): zero pivot in row " + std::to_string(b1));
        // the smallest pivot (a minimum: any order gives the same value)
        DevBuf mp(sizeof(double));
        launch_min_pivot(n, M->diag.as<int>(), M->lu.as<double>(), mp.as<double>(), s);
        double h = 0.0;
        HIPC(hipMemcpyAsync(&h, mp.p, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        M->min_pivot = n ? h : std::numeric_limits<double>::infinity();
    }
    plan_sweeps(M, pip, pix, perm, s);
    HIPC(hipStreamSynchronize(s));
}

thread_local hipEvent_t t_ev0 = nullptr, t_ev1 = nullptr;

} // namespace femd_amd

// ================================================================================== C ABI
extern "C" {

const char *femd_last_error(void) { return g_err.c_str(); }

int femd_init(int device)
{
    return api([&] {
        std::lock_guard<std::mutex> lk(g_init_mu);
        if (g_dev.ready) return;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) fail(FEMD_EDEVICE, "no HIP device is visible");
        if (device < 0 || device >= n) fail(FEMD_EINVAL, "device index out of range");
        g_dev.id = device;
        HIPC(hipSetDevice(device));
        HIPC(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
        noalloc_setup(device);
        g_dev.ready = true;
    });
}

int femd_dev_alloc(size_t bytes, void **out)
{
    return api([&] {
        ensure_device();
        void *p = nullptr;
        HIPC(hipMalloc(&p, std::max<size_t>(bytes, 1)));
        *out = p;
    });
}
int femd_dev_free(void *p) { return api([&] { if (p) HIPC(hipFree(p)); }); }
// Copies on the library stream, after the work already queued there, then synchronized.
int femd_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    return api([&] {
        hipStream_t s = S(nullptr);
        if (bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        HIPC(hipStreamSynchronize(s));
    });
}
int femd_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    return api([&] {
        hipStream_t s = S(nullptr);
        if (bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}
int femd_stream_sync(femd_stream s) { return api([&] { HIPC(hipStreamSynchronize(S(s))); }); }
int femd_vec_copy_dev(int64_t n, const double *src, double *dst, femd_stream s)
{
    return api([&] { launch_copy(n, src, dst, S(s)); });
}

// ---- meshes ------------------------------------------------------------------------------
int femd_mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, int diag, const int markers[4], femd_mesh *out)
{
    return api([&] {
        const int dm[4] = {1, 2, 3, 4};
        auto m = std::make_unique<femd_mesh_s>();
        m->M = mesh_rectangle(x0, y0, x1, y1, nx, ny, diag, markers ? markers : dm);
        *out = m.release();
    });
}

int femd_mesh_from_arrays(int np, const double *points, int nt, const int *triangles, int ns, const int *segments,
                          const int *smarks, int marker, femd_mesh *out)
{
    return api([&] {
        if (np < 0 || nt < 0 || ns < 0) fail(FEMD_EINVAL, "mesh_from_arrays: negative count");
        auto m = std::make_unique<femd_mesh_s>();
        m->M = mesh_from_arrays(std::vector<double>(points, points + 2 * size_t(np)), std::vector<int>(triangles, triangles + 3 * size_t(nt)),
                                std::vector<int>(segments, segments + 2 * size_t(ns)),
                                smarks ? std::vector<int>(smarks, smarks + ns) : std::vector<int>{}, marker);
        *out = m.release();
    });
}

int femd_quad_mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, const int markers[4], femd_mesh *out)
{
    return api([&] {
        const int dm[4] = {1, 2, 3, 4};
        auto m = std::make_unique<femd_mesh_s>();
        m->M = quad_mesh_rectangle(x0, y0, x1, y1, nx, ny, markers ? markers : dm);
        *out = m.release();
    });
}

int femd_quad_mesh_from_arrays(int np, const double *points, int nq, const int *quads, int ns, const int *segments,
                               const int *smarks, int marker, femd_mesh *out)
{
    return api([&] {
        if (np < 0 || nq < 0 || ns < 0) fail(FEMD_EINVAL, "quad_mesh_from_arrays: negative count");
        auto m = std::make_unique<femd_mesh_s>();
        m->M = quad_mesh_from_arrays(std::vector<double>(points, points + 2 * size_t(np)), std::vector<int>(quads, quads + 4 * size_t(nq)),
                                     std::vector<int>(segments, segments + 2 * size_t(ns)),
                                     smarks ? std::vector<int>(smarks, smarks + ns) : std::vector<int>{}, marker);
        *out = m.release();
    });
}

int femd_mesh_cell_vertices(femd_mesh m, int *nv)
{
    return api([&] {
        if (!m || !nv) fail(FEMD_EINVAL, "null argument");
        *nv = m->M.nv;
    });
}

int femd_mesh_sizes(femd_mesh m, int out[3])
{
    return api([&] {
        if (!m) fail(FEMD_EINVAL, "null mesh");
        out[0] = m->M.np(); out[1] = m->M.nt(); out[2] = static_cast<int>(m->M.smark.size());
    });
}

int femd_mesh_arrays(femd_mesh m, double *points, int *triangles, int *neighbours, int *segments, int *smarks)
{
    return api([&] {
        if (!m) fail(FEMD_EINVAL, "null mesh");
        const MeshData &M = m->M;
        if (points) std::copy(M.xy.begin(), M.xy.end(), points);
        if (triangles) std::copy(M.tri.begin(), M.tri.end(), triangles);
        if (neighbours) std::copy(M.nbr.begin(), M.nbr.end(), neighbours);
        if (segments) std::copy(M.seg.begin(), M.seg.end(), segments);
        if (smarks) std::copy(M.smark.begin(), M.smark.end(), smarks);
    });
}

void femd_mesh_free(femd_mesh m) { delete m; }

// ---- spaces ------------------------------------------------------------------------------
int femd_space_create(femd_mesh m, int degree, int ndirichlet, const int *dirichlet, int all_exterior, int equispaced,
                      femd_space *out)
{
    return api([&] {
        if (!m) fail(FEMD_EINVAL, "null mesh");
        if (degree < 1) fail(FEMD_EINVAL, "LagrangeSpace2D: degree must be >= 1 (P0 is a DG space)");
        const bool quad = m->M.nv == 4;
        if (!quad && degree > 5) fail(FEMD_EINVAL, "LagrangeSpace2D: the GPU kernels cover degrees 1..5");
        if (quad && degree > 4) fail(FEMD_EINVAL, "LagrangeSpaceQ: the GPU kernels cover degrees 1..4 on quadrilaterals");
        auto V = std::make_unique<femd_space_s>();
        V->M = m->M;
        V->V = build_space(m->M, degree, std::vector<int>(dirichlet, dirichlet + std::max(ndirichlet, 0)), all_exterior != 0,
                           equispaced != 0);
        if (quad) V->refq = std::make_unique<RefQuad>(degree, equispaced != 0);
        else V->ref = std::make_unique<RefElement>(degree, equispaced != 0);
        *out = V.release();
    });
}

int femd_space_sizes(femd_space V, int out[6])
{
    return api([&] {
        const SpaceData &D = check_space(V)->V;
        out[0] = D.ncells; out[1] = D.nloc; out[2] = D.raw_dim; out[3] = D.dim; out[4] = D.nedges; out[5] = D.npoints;
    });
}

int femd_space_numbering(femd_space V, int *cell_dof, int *cell_raw, int *ad2raw, double *raw_xy)
{
    return api([&] {
        const SpaceData &D = check_space(V)->V;
        if (cell_dof) std::copy(D.cell_dof.begin(), D.cell_dof.end(), cell_dof);
        if (cell_raw) std::copy(D.cell_raw.begin(), D.cell_raw.end(), cell_raw);
        if (ad2raw) std::copy(D.ad2raw.begin(), D.ad2raw.end(), ad2raw);
        if (raw_xy) std::copy(D.raw_xy.begin(), D.raw_xy.end(), raw_xy);
    });
}

void femd_space_free(femd_space V)
{
    if (V && g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
    delete V;
}

int femd_pattern(femd_space V, int64_t *nnz, int *indptr, int *indices)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        ensure_host_pattern(Vs);
        if (nnz) *nnz = Vs->P.nnz();
        if (indptr) std::copy(Vs->P.indptr.begin(), Vs->P.indptr.end(), indptr);
        if (indices) std::copy(Vs->P.indices.begin(), Vs->P.indices.end(), indices);
    });
}

int femd_pattern_lists(femd_space V, int where, int64_t *nnz, int64_t *ncon, int *indptr, int *indices, int *mptr,
                       int *mlist, int *vptr, int *vlist)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        if (where != 0 && where != 1) fail(FEMD_EINVAL, "pattern_lists: where is 0 (host) or 1 (GPU)");
        const int n = Vs->V.dim;
        auto put = [](const int *src, size_t k, int *dst) { if (dst && k) std::copy(src, src + k, dst); };
        if (where == 0)
        {
            const PatternData P = build_pattern(Vs->V);
            if (nnz) *nnz = P.nnz();
            if (ncon) *ncon = static_cast<int64_t>(P.mlist.size());
            put(P.indptr.data(), P.indptr.size(), indptr);
            put(P.indices.data(), P.indices.size(), indices);
            put(P.mptr.data(), P.mptr.size(), mptr);
            put(P.mlist.data(), P.mlist.size(), mlist);
            put(P.vptr.data(), P.vptr.size(), vptr);
            put(P.vlist.data(), P.vlist.size(), vlist);
            return;
        }
        hipStream_t s = S(nullptr);
        DevBuf cd = upload(Vs->V.cell_dof, s);
        DevPatternLists D;
        device_build_pattern(n, Vs->V.ncells, Vs->V.nloc, cd.as<int>(), D, s);
        if (nnz) *nnz = D.nnz;
        if (ncon) *ncon = D.ncon;
        put(D.h_ip.data(), D.h_ip.size(), indptr);
        put(D.h_ix.data(), D.h_ix.size(), indices);
        auto get = [&](const DevBuf &b, size_t k, int *dst) {
            if (dst && k) HIPC(hipMemcpyAsync(dst, b.p, 4 * k, hipMemcpyDeviceToHost, s));
        };
        int64_t nv = 0;
        for (int x : Vs->V.cell_dof) nv += x >= 0;
        get(D.mptr, size_t(D.nnz) + 1, mptr);
        get(D.mlist, size_t(D.ncon), mlist);
        get(D.vptr, size_t(n) + 1, vptr);
        get(D.vlist, size_t(nv), vlist);
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_program_compact(const femd_coefficient *c, int *code_out, int *nreg_out)
{
    return api([&] {
        check_program(*c);
        std::vector<int> code = compact_registers(*c, *nreg_out);
        std::copy(code.begin(), code.end(), code_out);
    });
}

int femd_quadrature_square(int degree, int *nq, double *xi, double *eta, double *w)
{
    return api([&] {
        QuadRule R = square_rule(degree);
        *nq = R.size();
        if (xi) std::copy(R.xi.begin(), R.xi.end(), xi);
        if (eta) std::copy(R.eta.begin(), R.eta.end(), eta);
        if (w) std::copy(R.w.begin(), R.w.end(), w);
    });
}

int femd_quadrature(int degree, int *nq, double *xi, double *eta, double *w)
{
    return api([&] {
        QuadRule R = triangle_rule(degree);
        *nq = R.size();
        if (xi) std::copy(R.xi.begin(), R.xi.end(), xi);
        if (eta) std::copy(R.eta.begin(), R.eta.end(), eta);
        if (w) std::copy(R.w.begin(), R.w.end(), w);
    });
}

// ---- CSR ---------------------------------------------------------------------------------
int femd_csr_create_space(femd_space V, femd_csr *out)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        hipStream_t s = S(nullptr);
        ensure_pattern(Vs, s);
        auto A = std::make_unique<femd_csr_s>();
        A->pat = Vs->pat;
        A->space = Vs;
        A->v = DevBuf(8 * size_t(std::max<int64_t>(A->pat->nnz, 1)));
        launch_fill(A->pat->nnz, 0.0, A->v.as<double>(), s);
        HIPC(hipStreamSynchronize(s));
        *out = A.release();
    });
}

int femd_csr_create(int nrows, int ncols, const int *indptr, const int *indices, const double *data, femd_csr *out)
{
    return api([&] {
        if (nrows < 0 || ncols < 0) fail(FEMD_EINVAL, "CSRMatrix: negative shape");
        if (indptr[0] != 0) fail(FEMD_EINVAL, "CSRMatrix: indptr[0] must be 0");
        for (int i = 0; i < nrows; ++i)
        {
            if (indptr[i + 1] < indptr[i]) fail(FEMD_EINVAL, "CSRMatrix: indptr must not decrease");
            for (int p = indptr[i]; p < indptr[i + 1]; ++p)
            {
                if (indices[p] < 0 || indices[p] >= ncols) fail(FEMD_EINVAL, "CSRMatrix: column index out of range");
                if (p > indptr[i] && indices[p] <= indices[p - 1]) fail(FEMD_EINVAL, "CSRMatrix: columns must be sorted and unique in each row");
            }
        }
        hipStream_t s = S(nullptr);
        const int64_t nnz = indptr[nrows];
        auto A = std::make_unique<femd_csr_s>();
        A->pat = make_pattern(nrows, ncols, std::vector<int>(indptr, indptr + nrows + 1), std::vector<int>(indices, indices + nnz), s);
        A->v = DevBuf(8 * size_t(std::max<int64_t>(nnz, 1)));
        if (nnz) HIPC(hipMemcpyAsync(A->v.p, data, 8 * size_t(nnz), hipMemcpyHostToDevice, s));
        HIPC(hipStreamSynchronize(s));
        *out = A.release();
    });
}

int femd_csr_info(femd_csr A, int out[2], int64_t *nnz)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        out[0] = a->pat->n; out[1] = a->pat->m;
        if (nnz) *nnz = a->pat->nnz;
    });
}

int femd_csr_download(femd_csr A, int *indptr, int *indices, double *data)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        if (indptr) std::copy(a->pat->h_ip.begin(), a->pat->h_ip.end(), indptr);
        if (indices) std::copy(a->pat->h_ix.begin(), a->pat->h_ix.end(), indices);
        if (data && a->pat->nnz)
        {
            hipStream_t s = S(nullptr);
            HIPC(hipMemcpyAsync(data, a->v.p, 8 * size_t(a->pat->nnz), hipMemcpyDeviceToHost, s));
            HIPC(hipStreamSynchronize(s));
        }
    });
}

int femd_csr_upload_values(femd_csr A, const double *data)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        if (a->pat->nnz) HIPC(hipMemcpyAsync(a->v.p, data, 8 * size_t(a->pat->nnz), hipMemcpyHostToDevice, s));
        a->v32_valid = false; a->v32s_valid = false; a->v64s_valid = false;
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_csr_device_arrays(femd_csr A, const int **indptr, const int **indices, double **data)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        if (indptr) *indptr = a->pat->ip.as<int>();
        if (indices) *indices = a->pat->ix.as<int>();
        if (data) *data = a->v.as<double>();
        a->v32_valid = false; a->v32s_valid = false; a->v64s_valid = false;                  // the caller may write the values
    });
}

int femd_csr_asymmetry(femd_csr A, double *out)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        if (a->pat->n != a->pat->m) { *out = 1.0; return; }
        hipStream_t s = S(nullptr);
        DevPattern &P = *a->pat;
        if (!P.have_tpos)
        {
            P.tpos = DevBuf(4 * size_t(std::max<int64_t>(P.nnz, 1)));
            launch_transpose_positions(P.n, P.ip.as<int>(), P.ix.as<int>(), P.tpos.as<int>(), s);
            P.have_tpos = true;
        }
        DevBuf r(16);
        launch_fill(2, 0.0, r.as<double>(), s);
        launch_asymmetry(P.nnz, P.tpos.as<int>(), a->v.as<double>(), r.as<double>(), s);
        double h[2];
        HIPC(hipMemcpyAsync(h, r.p, 16, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        *out = h[0] > 0.0 ? h[1] / h[0] : 0.0;
    });
}

int femd_csr_diagonal(femd_csr A, double *out)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        const int n = std::min(a->pat->n, a->pat->m);
        const int *dg = diag_of(*a->pat, s);
        std::vector<int> d(n);
        std::vector<double> v(a->pat->nnz);
        HIPC(hipMemcpyAsync(d.data(), dg, 4 * size_t(n), hipMemcpyDeviceToHost, s));
        if (a->pat->nnz) HIPC(hipMemcpyAsync(v.data(), a->v.p, 8 * size_t(a->pat->nnz), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) out[i] = d[i] >= 0 ? v[d[i]] : 0.0;
    });
}

void femd_csr_free(femd_csr A)
{
    if (A && g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
    delete A;
}

// ---- assembly ----------------------------------------------------------------------------
extern "C++" {
namespace {
std::atomic<int> g_assembly_mode{0};       // 0: exact VALU element kernels, 1: FP64 matrix cores

// The plan of the matrix-core element kernel for one dx group (see mfma.hip): the reference-row
// pairs (alpha, beta) the terms reach -- the value row for code 0, the two derivative rows for
// codes 1 and 2 -- merged with their transposes when the terms are closed under swapping their
// codes with the same coefficient; R[c, q][n] from the reference tables at the rule's points.
// nullptr when the group is outside the kernel's sizes (the exact kernel runs instead).
// prec: 0 FP64 (steps of four k), 1 f16, 2 bf16 (steps of sixteen k, R rounded to 16 bits)
const MfmaPlan *mfma_plan(femd_space_s *V, int degree, const Quad &Q, int nterms, const int *a, const int *b,
                          const femd_coefficient *coeffs, hipStream_t s, int prec = 0)
{
    bool sym = true;
    for (int t = 0; t < nterms && sym; ++t)
    {
        if (a[t] == b[t]) continue;
        bool found = false;
        for (int u = 0; u < nterms && !found; ++u)
            found = a[u] == b[t] && b[u] == a[t] && same_coefficient(coeffs[t], coeffs[u]);
        sym = found;
    }
    std::string key = std::to_string(prec) + ":" + std::to_string(degree) + (sym ? "s" : "n");
    for (int t = 0; t < nterms; ++t) key += ":" + std::to_string(a[t]) + std::to_string(b[t]);
    auto it = V->mfma_plans.find(key);
    if (it != V->mfma_plans.end()) return it->second->rt.p ? it->second.get() : nullptr;

    const int N = V->V.nloc, nq = Q.nq;
    bool reach[3][3] = {};
    for (int t = 0; t < nterms; ++t)
        for (int al = 0; al < 3; ++al)
            for (int be = 0; be < 3; ++be)
                if ((a[t] == 0) == (al == 0) && (b[t] == 0) == (be == 0)) reach[al][be] = true;
    std::vector<std::pair<int, int>> combos;
    for (int al = 0; al < 3; ++al)
        for (int be = 0; be < 3; ++be)
            if (reach[al][be] && (!sym || al <= be)) combos.push_back({al, be});
    auto P = std::make_unique<MfmaPlan>();
    P->ncombo = static_cast<int>(combos.size());
    const int kstep = prec ? 16 : 4;
    P->spc = (nq + kstep - 1) / kstep;         // each combo's points padded to whole steps
    P->ks = P->ncombo * P->spc;
    P->nout = sym ? N * (N + 1) / 2 : N * N;
    const int tiles = (P->nout + 15) / 16;
    P->nt = std::min(tiles, 4);
    if (prec && P->nt == 3) P->nt = 4;
    P->ntile = (tiles + P->nt - 1) / P->nt * P->nt;
    for (int c = 0; c < P->ncombo; ++c)
    {
        P->alpha |= unsigned(combos[c].first) << (2 * c);
        P->beta |= unsigned(combos[c].second) << (2 * c);
    }
    const MfmaPlan *out = nullptr;
    if (P->spc <= (prec ? 2 : MFMA_SPCMAX))
    {
        // the output columns: (i, j) for i <= j (symmetric) or every (i, j), row-major
        std::vector<int> oc(2 * size_t(16) * P->ntile, -1), ii, jj;
        for (int i = 0; i < N; ++i)
            for (int j = sym ? i : 0; j < N; ++j) { ii.push_back(i); jj.push_back(j); }
        for (int n = 0; n < P->nout; ++n)
        {
            oc[2 * n] = ii[n] * N + jj[n];
            oc[2 * n + 1] = sym && ii[n] != jj[n] ? jj[n] * N + ii[n] : -1;
        }
        const int npad = 16 * P->ntile;
        std::vector<double> rt(size_t(kstep) * P->ks * npad, 0.0), tab(size_t(3) * N);
        for (int q = 0; q < nq; ++q)
        {
            if (V->V.nv == 4) V->refq->tabulate(Q.xi[q], Q.eta[q], tab.data());
            else V->ref->tabulate(Q.xi[q], Q.eta[q], tab.data());
            for (int c = 0; c < P->ncombo; ++c)
            {
                const int al = combos[c].first, be = combos[c].second;
                double *row = &rt[size_t(c * kstep * P->spc + q) * npad];
                for (int n = 0; n < P->nout; ++n)
                {
                    const int i = ii[n], j = jj[n];
                    double r = tab[al * N + i] * tab[be * N + j];
                    if (sym && al != be) r += tab[be * N + i] * tab[al * N + j];
                    row[n] = r;
                }
            }
        }
        if (prec)
        {
            // 16-bit R laid out [step][entry][16 k], rounded to nearest even
            std::vector<uint16_t> r16(size_t(16) * P->ks * npad, 0);
            for (int st = 0; st < P->ks; ++st)
                for (int n = 0; n < npad; ++n)
                    for (int k = 0; k < 16; ++k)
                    {
                        const float f = static_cast<float>(rt[size_t(st * 16 + k) * npad + n]);
                        uint16_t h;
                        if (prec == 1)
                        {
                            const _Float16 x = static_cast<_Float16>(f);
                            std::memcpy(&h, &x, 2);
                        }
                        else
                        {
                            uint32_t u;
                            std::memcpy(&u, &f, 4);
                            h = static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
                        }
                        r16[(size_t(st) * npad + n) * 16 + k] = h;
                    }
            P->rlds = false;
            P->rt = upload(r16, s);
        }
        else
        {
            P->rlds = mfma_table_bytes(P->ks, P->ntile) <= MFMA_LDS_MAX;
            P->rt = upload(rt, s);
        }
        P->ocol = upload(oc, s);
        HIPC(hipStreamSynchronize(s));
        out = P.get();
    }
    V->mfma_plans[key] = std::move(P);
    return out;
}

MfmaJob mfma_job(femd_space_s *V, Quad &Q, const MfmaPlan &P, double *E)
{
    return MfmaJob{V->V.ncells, Q.nq, V->V.nloc, P.ncombo, P.spc, P.ks, P.ntile, P.nout, P.alpha, P.beta,
                   V->geom.as<double>(), Q.qw.as<double>(), Q.qxi.as<double>(), Q.qeta.as<double>(),
                   P.rt.as<double>(), P.ocol.as<int2>(), E};
}

void assemble_matrix_dx(femd_space V, int quad_degree, int nterms, const int *a, const int *b, const femd_coefficient *coeffs,
                        femd_csr K, int mode)
{
    {
        femd_space_s *Vs = check_space(V);
        femd_csr_s *Ks = check_csr(K);
        hipStream_t s = S(nullptr);
        ensure_pattern(Vs, s);
        if (Ks->pat != Vs->pat) fail(FEMD_EPATTERN, "assemble_matrix: the matrix is not on the space's pattern");
        Quad &Q = quad_of(Vs, quad_degree, s);
        Coeffs C = make_coeffs(Vs, Q, quad_degree, nterms, a, b, coeffs, s);
        const int n = Vs->V.nloc, nc = Vs->V.ncells;
        DevBuf E(8 * std::max<size_t>(size_t(nc) * n * n, 1));
        const int amode = g_assembly_mode.load();
        const MfmaPlan *mp = amode >= 1 ? mfma_plan(Vs, quad_degree, Q, nterms, a, b, coeffs, s, amode - 1) : nullptr;
        bool done = false;
        if (mp && amode == 1)
            done = launch_elem_matrix_mfma(Vs->V.nv, mfma_job(Vs, Q, *mp, E.as<double>()), C.T, mp->nt, mp->rlds, s);
        else if (mp)
            done = launch_elem_matrix_mfma16(Vs->V.nv, mfma_job(Vs, Q, *mp, E.as<double>()), C.T, mp->nt, amode == 3, s);
        if (!done &&
            !launch_elem_matrix(Vs->V.nv, n, nc, Q.nq, Vs->geom.as<double>(), Q.tab.as<double>(), Q.qw.as<double>(),
                                Q.qxi.as<double>(), Q.qeta.as<double>(), C.T, E.as<double>(), s))
            fail(FEMD_EINVAL, "assemble_matrix: quadrature degree " + std::to_string(quad_degree) + " is too high for the element kernel");
        launch_gather_mode(Ks->pat->nnz, nullptr, Vs->mptr.as<int>(), Vs->mlist.as<int>(), E.as<double>(), Ks->v.as<double>(),
                           mode, s);
        Ks->v32_valid = false; Ks->v32s_valid = false; Ks->v64s_valid = false;
        HIPC(hipStreamSynchronize(s));
    }
}

void assemble_vector_dx(femd_space V, int quad_degree, int nterms, const int *a, const femd_coefficient *coeffs, double *f_dev,
                        hipStream_t s, int mode)
{
    femd_space_s *Vs = check_space(V);
    ensure_pattern(Vs, s);
    Quad &Q = quad_of(Vs, quad_degree, s);
    Coeffs C = make_coeffs(Vs, Q, quad_degree, nterms, a, nullptr, coeffs, s);
    const int n = Vs->V.nloc, nc = Vs->V.ncells;
    DevBuf F(8 * std::max<size_t>(size_t(nc) * n, 1));
    if (!launch_elem_vector(Vs->V.nv, n, nc, Q.nq, Vs->geom.as<double>(), Q.tab.as<double>(), Q.qw.as<double>(),
                            Q.qxi.as<double>(), Q.qeta.as<double>(), C.T, F.as<double>(), s))
        fail(FEMD_EINVAL, "assemble_vector: quadrature degree " + std::to_string(quad_degree) + " is too high for the element kernel");
    launch_gather_mode(Vs->V.dim, nullptr, Vs->vptr.as<int>(), Vs->vlist.as<int>(), F.as<double>(), f_dev, mode, s);
    HIPC(hipStreamSynchronize(s));
}
} // namespace
} // extern "C++"

int femd_assembly_mode(int mode)
{
    const int prev = g_assembly_mode.load();
    if (mode >= 0 && mode <= 3) g_assembly_mode.store(mode);
    return prev;
}

int femd_assemble_matrix(femd_space V, int quad_degree, int nterms, const int *a, const int *b, const femd_coefficient *coeffs,
                         femd_csr K)
{
    return api([&] { assemble_matrix_dx(V, quad_degree, nterms, a, b, coeffs, K, GATHER_SET); });
}

int femd_assemble_matrix_add(femd_space V, int quad_degree, int nterms, const int *a, const int *b,
                             const femd_coefficient *coeffs, femd_csr K)
{
    return api([&] { assemble_matrix_dx(V, quad_degree, nterms, a, b, coeffs, K, GATHER_CONTINUE); });
}

int femd_assemble_vector_dev(femd_space V, int quad_degree, int nterms, const int *a, const femd_coefficient *coeffs,
                             double *f_dev, femd_stream st)
{
    return api([&] { assemble_vector_dx(V, quad_degree, nterms, a, coeffs, f_dev, S(st), GATHER_SET_PZ); });
}

int femd_assemble_vector_add(femd_space V, int quad_degree, int nterms, const int *a, const femd_coefficient *coeffs, double *f)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        hipStream_t s = S(nullptr);
        const int dim = Vs->V.dim;
        DevBuf fd(8 * size_t(std::max(dim, 1)));
        if (dim) HIPC(hipMemcpyAsync(fd.p, f, 8 * size_t(dim), hipMemcpyHostToDevice, s));
        assemble_vector_dx(V, quad_degree, nterms, a, coeffs, fd.as<double>(), s, GATHER_ADD);
        if (dim) HIPC(hipMemcpyAsync(f, fd.p, 8 * size_t(dim), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_assemble_vector(femd_space V, int quad_degree, int nterms, const int *a, const femd_coefficient *coeffs, double *f)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        const int dim = Vs->V.dim;
        DevBuf fd(8 * size_t(std::max(dim, 1)));
        const int rc = femd_assemble_vector_dev(V, quad_degree, nterms, a, coeffs, fd.as<double>(), nullptr);
        if (rc) throw Error(rc, g_err);
        hipStream_t s = S(nullptr);
        if (dim) HIPC(hipMemcpyAsync(f, fd.p, 8 * size_t(dim), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_eval_coefficient(femd_space V, int quad_degree, const femd_coefficient *c, double *out)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        hipStream_t s = S(nullptr);
        ensure_geom(Vs, s);
        Quad &Q = quad_of(Vs, quad_degree, s);
        const size_t np = size_t(Vs->V.ncells) * Q.nq;
        DevBuf d(8 * std::max<size_t>(np, 1));
        eval_program(Vs, Q, quad_degree, *c, d.as<double>(), s);
        if (np) HIPC(hipMemcpyAsync(out, d.p, 8 * np, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

// ---- boundary facets (ds) -----------------------------------------------------------------
extern "C++" {
namespace {
femd_facets_s *check_facets(femd_facets F)
{
    if (!F || !F->space) fail(FEMD_EINVAL, "null facet set");
    return F;
}

template <class T> DevBuf upload_nz(const std::vector<T> &v, hipStream_t s)
{
    return v.empty() ? DevBuf(sizeof(T)) : upload(v, s);
}

// The facet set's arrays on the device, uploaded on first use (creation is host work only).
void facets_device(femd_facets_s *Fs, hipStream_t s)
{
    if (Fs->var.p) return;
    const FacetData &D = Fs->F;
    Fs->var = upload_nz(D.var, s); Fs->w = upload_nz(D.w, s); Fs->Ji = upload_nz(D.Ji, s); Fs->tab = upload_nz(D.tab, s);
    Fs->x = upload_nz(D.x, s); Fs->y = upload_nz(D.y, s); Fs->nx = upload_nz(D.nx, s); Fs->ny = upload_nz(D.ny, s);
    Fs->mpos = upload_nz(D.mpos, s); Fs->mptr = upload_nz(D.mptr, s); Fs->mlist = upload_nz(D.mlist, s);
    Fs->vdof = upload_nz(D.vdof, s); Fs->vptr = upload_nz(D.vptr, s); Fs->vlist = upload_nz(D.vlist, s);
    HIPC(hipStreamSynchronize(s));
}

// A program at the facet points, the outward normal available (Integrand NX / NY)
void eval_program_facets(femd_facets_s *Fs, const femd_coefficient &c, double *out, hipStream_t s)
{
    check_program(c, true);
    int nreg = 0;
    std::vector<int> compact = compact_registers(c, nreg);
    if (nreg > MAX_REGS) fail(FEMD_EINVAL, "coefficient: the program needs more than " + std::to_string(MAX_REGS) + " live registers");
    DevBuf code = upload(compact, s);
    DevBuf consts = upload(c.consts, static_cast<size_t>(std::max(c.nconst, 0)), s);
    launch_coefficient_pts(int64_t(Fs->F.nent) * Fs->F.nq, Fs->x.as<double>(), Fs->y.as<double>(), Fs->nx.as<double>(),
                           Fs->ny.as<double>(), c.ninstr, code.as<int>(), consts.as<double>(), out, s);
    HIPC(hipStreamSynchronize(s));
}

Coeffs facet_coeffs(femd_facets_s *Fs, int nt, const int *a, const int *b, const femd_coefficient *c, hipStream_t s)
{
    if (nt < 1 || nt > MAX_TERMS) fail(FEMD_EINVAL, "assembly: 1.." + std::to_string(MAX_TERMS) + " terms per call");
    Coeffs C;
    C.T.nt = nt;
    const size_t np = size_t(Fs->F.nent) * Fs->F.nq;
    for (int t = 0; t < nt; ++t)
    {
        if (a[t] < 0 || a[t] > 2 || (b && (b[t] < 0 || b[t] > 2))) fail(FEMD_EINVAL, "assembly_2d: derivative code is not in 0..2");
        C.T.a[t] = a[t];
        C.T.b[t] = b ? b[t] : 0;
        C.T.c[t] = nullptr;
        C.T.cval[t] = 0.0;
        const femd_coefficient &k = c[t];
        if (k.values)
        {
            C.bufs.push_back(np ? upload(k.values, np, s) : DevBuf(sizeof(double)));
            C.T.c[t] = C.bufs.back().as<double>();
        }
        else if (k.ninstr == 1 && k.code[0] == 0)
        {
            if (k.code[2] < 0 || k.code[2] >= k.nconst) fail(FEMD_EINVAL, "coefficient: constant out of range");
            C.T.cval[t] = k.consts[k.code[2]];
        }
        else
        {
            C.bufs.emplace_back(std::max<size_t>(np, 1) * sizeof(double));
            if (np) eval_program_facets(Fs, k, C.bufs.back().as<double>(), s);
            C.T.c[t] = C.bufs.back().as<double>();
        }
    }
    return C;
}
} // namespace
} // extern "C++"

int femd_facets_create(femd_space V, int degree, int nmarkers, const int *markers, int all, femd_facets *out)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        if (!all && (nmarkers < 1 || !markers)) fail(FEMD_EINVAL, "ds: name the markers, or ask for every exterior facet");
        ensure_host_pattern(Vs);
        auto F = std::make_unique<femd_facets_s>();
        F->space = Vs;
        F->F = boundary_facets(Vs->M, Vs->V, Vs->P, Vs->ref.get(), Vs->refq.get(), std::max(degree, 0),
                               all ? std::vector<int>{} : std::vector<int>(markers, markers + nmarkers), all != 0);
        *out = F.release();
    });
}

int femd_facets_sizes(femd_facets F, int out[2])
{
    return api([&] {
        femd_facets_s *Fs = check_facets(F);
        out[0] = Fs->F.nent; out[1] = Fs->F.nq;
    });
}

int femd_facets_points(femd_facets F, double *x, double *y, double *w, double *nx, double *ny)
{
    return api([&] {
        const FacetData &D = check_facets(F)->F;
        if (x) std::copy(D.x.begin(), D.x.end(), x);
        if (y) std::copy(D.y.begin(), D.y.end(), y);
        if (w) std::copy(D.w.begin(), D.w.end(), w);
        if (nx) std::copy(D.nx.begin(), D.nx.end(), nx);
        if (ny) std::copy(D.ny.begin(), D.ny.end(), ny);
    });
}

void femd_facets_free(femd_facets F)
{
    if (F && F->var.p && g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
    delete F;
}

int femd_assemble_matrix_facets(femd_facets F, int nterms, const int *a, const int *b, const femd_coefficient *coeffs,
                                femd_csr K, int mode)
{
    return api([&] {
        femd_facets_s *Fs = check_facets(F);
        femd_space_s *Vs = Fs->space;
        femd_csr_s *Ks = check_csr(K);
        hipStream_t s = S(nullptr);
        ensure_pattern(Vs, s);
        if (Ks->pat != Vs->pat) fail(FEMD_EPATTERN, "assemble_matrix: the matrix is not on the space's pattern");
        if (mode != 0 && mode != 1) fail(FEMD_EINVAL, "assemble_matrix_facets: mode is 0 (zero K first) or 1 (add)");
        facets_device(Fs, s);
        const FacetData &D = Fs->F;
        Coeffs C = facet_coeffs(Fs, nterms, a, b, coeffs, s);
        if (mode == 0) launch_fill(Ks->pat->nnz, 0.0, Ks->v.as<double>(), s);
        const int n = D.nloc;
        DevBuf E(8 * std::max<size_t>(size_t(D.nent) * n * n, 1));
        if (!launch_facet_matrix(n, D.nent, D.nq, Fs->var.as<int>(), Fs->w.as<double>(), Fs->Ji.as<double>(), Fs->tab.as<double>(),
                                 C.T, E.as<double>(), s))
            fail(FEMD_EINVAL, "assemble_matrix_facets: no facet kernel for this element");
        launch_gather_mode(static_cast<int64_t>(D.mpos.size()), Fs->mpos.as<int>(), Fs->mptr.as<int>(), Fs->mlist.as<int>(),
                           E.as<double>(), Ks->v.as<double>(), GATHER_CONTINUE, s);
        Ks->v32_valid = false; Ks->v32s_valid = false; Ks->v64s_valid = false;
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_assemble_vector_facets(femd_facets F, int nterms, const int *a, const femd_coefficient *coeffs, double *f, int mode)
{
    return api([&] {
        femd_facets_s *Fs = check_facets(F);
        femd_space_s *Vs = Fs->space;
        hipStream_t s = S(nullptr);
        if (mode != 0 && mode != 1) fail(FEMD_EINVAL, "assemble_vector_facets: mode is 0 (zero f first) or 1 (add)");
        facets_device(Fs, s);
        const FacetData &D = Fs->F;
        const int dim = Vs->V.dim, n = D.nloc;
        Coeffs C = facet_coeffs(Fs, nterms, a, nullptr, coeffs, s);
        DevBuf fd(8 * size_t(std::max(dim, 1)));
        if (mode == 0) launch_fill(dim, 0.0, fd.as<double>(), s);
        else if (dim) HIPC(hipMemcpyAsync(fd.p, f, 8 * size_t(dim), hipMemcpyHostToDevice, s));
        DevBuf Fe(8 * std::max<size_t>(size_t(D.nent) * n, 1));
        if (!launch_facet_vector(n, D.nent, D.nq, Fs->var.as<int>(), Fs->w.as<double>(), Fs->Ji.as<double>(), Fs->tab.as<double>(),
                                 C.T, Fe.as<double>(), s))
            fail(FEMD_EINVAL, "assemble_vector_facets: no facet kernel for this element");
        launch_gather_mode(static_cast<int64_t>(D.vdof.size()), Fs->vdof.as<int>(), Fs->vptr.as<int>(), Fs->vlist.as<int>(),
                           Fe.as<double>(), fd.as<double>(), GATHER_ADD, s);
        if (dim) HIPC(hipMemcpyAsync(f, fd.p, 8 * size_t(dim), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_eval_coefficient_facets(femd_facets F, const femd_coefficient *c, double *out)
{
    return api([&] {
        femd_facets_s *Fs = check_facets(F);
        hipStream_t s = S(nullptr);
        facets_device(Fs, s);
        const size_t np = size_t(Fs->F.nent) * Fs->F.nq;
        DevBuf d(8 * std::max<size_t>(np, 1));
        if (np) eval_program_facets(Fs, *c, d.as<double>(), s);
        if (np) HIPC(hipMemcpyAsync(out, d.p, 8 * np, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

// ---- SpMV and dot --------------------------------------------------------------------------
int femd_spmv_dev(femd_csr A, const void *x, void *y, int precision, femd_stream st)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(st);
        DevPattern &P = *a->pat;
        if (precision == FEMD_FP32)
        {
            const float *v32 = values32(a, s);
            spmv_run32(spmv_plan(P, FEMD_FP32, v32, s), P, v32, static_cast<const float *>(x), static_cast<float *>(y),
                       nullptr, s);
        }
        else spmv64(a, static_cast<const double *>(x), static_cast<double *>(y), nullptr, s);
    });
}

int femd_spmv(femd_csr A, const double *x, double *y)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        DevBuf xd = upload(x, size_t(a->pat->m), s), yd(8 * size_t(std::max(a->pat->n, 1)));
        spmv64(a, xd.as<double>(), yd.as<double>(), nullptr, s);
        if (a->pat->n) HIPC(hipMemcpyAsync(y, yd.p, 8 * size_t(a->pat->n), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_ddot_dev(int64_t n, const double *x, const double *y, double *out, femd_stream st)
{
    return api([&] {
        hipStream_t s = S(st);
        DevBuf parts(8 * size_t(std::max(nchunks(n), 1))), r(8);
        ddot_dev(n, x, y, parts.as<double>(), r.as<double>(), s);
        *out = read_scalar(r.as<double>(), s);
    });
}

int femd_ddot(int64_t n, const double *x, const double *y, double *out)
{
    return api([&] {
        hipStream_t s = S(nullptr);
        DevBuf xd = upload(x, size_t(n), s), yd = upload(y, size_t(n), s);
        DevBuf parts(8 * size_t(std::max(nchunks(n), 1))), r(8);
        ddot_dev(n, xd.as<double>(), yd.as<double>(), parts.as<double>(), r.as<double>(), s);
        *out = read_scalar(r.as<double>(), s);
    });
}

// ---- preconditioners -----------------------------------------------------------------------
int femd_precond_create(femd_csr A, int kind, int ordering, double omega, femd_precond *out, double *min_pivot)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        if (a->pat->n != a->pat->m) fail(FEMD_EINVAL, "a preconditioner needs a square matrix");
        if (kind < FEMD_PC_JACOBI || kind > FEMD_PC_SSOR) fail(FEMD_EINVAL, "preconditioner kind must be Jacobi, ILU(0) or SSOR");
        if (ordering != FEMD_ORDER_NATURAL && ordering != FEMD_ORDER_RCM) fail(FEMD_EINVAL, "ordering is natural or rcm");
        if (kind == FEMD_PC_SSOR && !(omega > 0.0 && omega < 2.0)) fail(FEMD_EINVAL, "SSOR: omega must lie in (0, 2)");
        auto M = std::make_unique<femd_precond_s>();
        M->kind = kind;
        M->ordering = kind == FEMD_PC_JACOBI ? FEMD_ORDER_NATURAL : ordering;
        M->omega = omega;
        build_precond(M.get(), a, S(nullptr));
        if (min_pivot) *min_pivot = M->min_pivot;
        *out = M.release();
    });
}

int femd_precond_apply_dev(femd_precond M, const double *r, double *z, femd_stream st)
{
    return api([&] {
        if (!M) fail(FEMD_EINVAL, "null preconditioner");
        precond_apply(M, r, z, nullptr, S(st));
    });
}

int femd_precond_apply(femd_precond M, const double *r, double *z)
{
    return api([&] {
        if (!M) fail(FEMD_EINVAL, "null preconditioner");
        hipStream_t s = S(nullptr);
        DevBuf rd = upload(r, size_t(M->n), s), zd(8 * size_t(std::max(M->n, 1)));
        precond_apply(M, rd.as<double>(), zd.as<double>(), nullptr, s);
        if (M->n) HIPC(hipMemcpyAsync(z, zd.p, 8 * size_t(M->n), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_precond_stats(femd_precond M, int64_t out[8])
{
    return api([&] {
        if (!M || !out) fail(FEMD_EINVAL, "precond_stats: null argument");
        out[0] = M->sweep;
        out[1] = M->nstepsF;
        out[2] = M->nstepsB;
        out[3] = M->nlevF;
        out[4] = M->nlevB;
        out[5] = M->bytesF;
        out[6] = M->bytesB;
        out[7] = M->entriesLU;
    });
}

int femd_precond_info(femd_precond M, int out[3])
{
    return api([&] {
        if (!M) fail(FEMD_EINVAL, "null preconditioner");
        out[0] = M->kind; out[1] = M->ordering; out[2] = M->n;
    });
}

int femd_inner_order_pattern(int n, const int *indptr, const int *indices, int nseg, const int64_t *segments, int *perm,
                             int *indptr2, int *indices2, int *emap)
{
    return api([&] {
        if (n < 0 || nseg < 0 || !indptr || (n > 0 && !indices) || !perm || !indptr2 || (nseg > 0 && !segments))
            fail(FEMD_EINVAL, "femd_inner_order_pattern: bad arguments");
        if (indptr[0] != 0) fail(FEMD_EINVAL, "femd_inner_order_pattern: indptr[0] must be 0");
        for (int i = 0; i < n; ++i)
            if (indptr[i + 1] < indptr[i]) fail(FEMD_EINVAL, "femd_inner_order_pattern: indptr decreases");
        if (n > 0 && indptr[n] > 0 && (!indices2 || !emap)) fail(FEMD_EINVAL, "femd_inner_order_pattern: null output");
        std::vector<int64_t> seg;
        if (nseg > 0)
        {
            seg.assign(segments, segments + nseg + 1);
            if (seg.front() != 0 || seg.back() != n) fail(FEMD_EINVAL, "femd_inner_order_pattern: segments must cover the rows");
            for (int k = 0; k < nseg; ++k)
                if (seg[size_t(k) + 1] < seg[size_t(k)]) fail(FEMD_EINVAL, "femd_inner_order_pattern: segments decrease");
        }
        for (int p = 0; p < (n > 0 ? indptr[n] : 0); ++p)
            if (indices[p] < 0 || indices[p] >= n) fail(FEMD_EINVAL, "femd_inner_order_pattern: column index out of range");
        std::vector<int> pm, ip2, ix2, em;
        if (!first_met_order(n, indptr, indices, seg, pm, ip2, ix2, em))
            fail(FEMD_EINVAL, "femd_inner_order_pattern: a segment's rows have columns outside it");
        std::copy(pm.begin(), pm.end(), perm);
        std::copy(ip2.begin(), ip2.end(), indptr2);
        std::copy(ix2.begin(), ix2.end(), indices2);
        std::copy(em.begin(), em.end(), emap);
    });
}

int femd_sweep_host_apply(int n, const int *indptr, const int *indices, const double *values, int kind, double omega,
                          const double *r, double *z, int stats[6])
{
    return api([&] {
        if (n < 0 || !indptr || (n > 0 && (!indices || !values || !r || !z)))
            fail(FEMD_EINVAL, "sweep_host_apply: null argument");
        if (kind != FEMD_PC_ILU0 && kind != FEMD_PC_SSOR) fail(FEMD_EINVAL, "sweep_host_apply: kind is ILU0 or SSOR");
        if (n == 0) return;
        std::vector<int> dg(n, -1);
        for (int i = 0; i < n; ++i)
            for (int p = indptr[i]; p < indptr[i + 1]; ++p)
                if (indices[p] == i) dg[i] = p;
        for (int i = 0; i < n; ++i)
            if (dg[i] < 0) fail(FEMD_EZEROPIVOT, "sweep_host_apply: no diagonal entry in row " + std::to_string(i));
        const bool ilu = kind == FEMD_PC_ILU0;
        SweepPlan B, F;
        std::string why;
        if (!plan_sweep(ilu ? SW_ILU_BWD : SW_SSOR_BWD, n, indptr, indices, dg.data(), SW_ROWS, B, why))
            fail(FEMD_EINVAL, "sweep_host_apply: " + why);
        if (!plan_sweep(ilu ? SW_ILU_FWD : SW_SSOR_FWD, n, indptr, indices, dg.data(), SW_ROWS, F, why))
            fail(FEMD_EINVAL, "sweep_host_apply: " + why);
        std::vector<uint8_t> fs = F.stream, bs = B.stream;
        double *f64 = reinterpret_cast<double *>(fs.data()), *b64 = reinterpret_cast<double *>(bs.data());
        for (int64_t p = 0; p < indptr[n]; ++p)
        {
            if (F.vmap[p] >= 0) f64[F.vmap[p]] = values[p];
            if (B.vmap[p] >= 0) b64[B.vmap[p]] = values[p];
        }
        for (int i = 0; i < n; ++i) f64[F.in_pos[i]] = r[i];
        std::vector<double> stage(n);
        if (!emulate_sweep(F, fs, omega, stage.data(), why)) fail(FEMD_EINVAL, "sweep_host_apply: " + why);
        for (int k = 0; k < n; ++k) b64[B.in_pos[F.order[k]]] = stage[k];
        if (!emulate_sweep(B, bs, omega, stage.data(), why)) fail(FEMD_EINVAL, "sweep_host_apply: " + why);
        for (int k = 0; k < n; ++k) z[B.order[k]] = stage[k];
        if (stats)
        {
            stats[0] = F.nsteps; stats[1] = B.nsteps; stats[2] = F.nlevels; stats[3] = B.nlevels;
            stats[4] = F.nslots; stats[5] = B.nslots;
        }
    });
}

int femd_rcm(int n, const int *indptr, const int *indices, int *perm)
{
    return api([&] {
        std::vector<int> p = rcm_order(n, indptr, indices);
        std::copy(p.begin(), p.end(), perm);
    });
}

void femd_precond_free(femd_precond M)
{
    if (M && g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
    delete M;
}

// ---- solvers --------------------------------------------------------------------------------
int femd_cg_dev(femd_csr A, femd_precond M, const double *b, double *x, int max_iter, double tol, femd_solve_info *info,
                femd_stream st)
{
    return api([&] { cg_device(check_csr(A), M, b, x, max_iter, tol, info, S(st)); });
}

int femd_cg_dev_noalloc(femd_csr A, femd_precond M, const double *b, double *x, int max_iter, double tol,
                        femd_solve_info *info, femd_stream st)
{
    return api([&] {
        if (!info) fail(FEMD_EINVAL, "cg_dev_noalloc: null info");
        cg_noalloc(check_csr(A), M, b, x, max_iter, tol, info, S(st));
    });
}

int femd_cg(femd_csr A, femd_precond M, const double *b, double *x, int max_iter, double tol, femd_solve_info *info)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        const int n = a->pat->n;
        DevBuf bd = upload(b, size_t(n), s), xd = upload(x, size_t(n), s);
        cg_device(a, M, bd.as<double>(), xd.as<double>(), max_iter, tol, info, s);
        if (n) HIPC(hipMemcpyAsync(x, xd.p, 8 * size_t(n), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_krylov(femd_csr A, femd_precond M, const double *b, double *x, int lgmres, int restart, int max_iter, double tol,
                int k_aug, femd_solve_info *info)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        if (a->pat->n != a->pat->m) fail(FEMD_EINVAL, "krylov: the matrix must be square");
        if (M && M->n != a->pat->n) fail(FEMD_EINVAL, "krylov: preconditioner and right-hand side differ in size");
        if (restart <= 0 || max_iter <= 0) fail(FEMD_EINVAL, "gmres: restart and max_iter must be positive");
        if (restart + (lgmres ? std::max(k_aug, 0) : 0) + 2 > 512) fail(FEMD_EINVAL, "gmres: restart (+ k_aug) must stay below 510");
        hipStream_t s = S(nullptr);
        const int n = a->pat->n;
        DevBuf bd = upload(b, size_t(n), s), xd = upload(x, size_t(n), s);
        krylov_device(a, M, bd.as<double>(), xd.as<double>(), lgmres, restart, max_iter, tol, k_aug, info, s);
        if (n) HIPC(hipMemcpyAsync(x, xd.p, 8 * size_t(n), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

int femd_solve_refined_dev(femd_csr A, const double *b, double *x, double tol, double inner_tol, int max_outer,
                           femd_solve_info *info, femd_stream st)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        if (a->pat->n != a->pat->m) fail(FEMD_EINVAL, "solve: the matrix must be square");
        refined_device(a, b, x, tol, inner_tol, max_outer, info, S(st));
    });
}

int femd_solve_refined(femd_csr A, const double *b, double *x, double tol, double inner_tol, int max_outer, femd_solve_info *info)
{
    return api([&] {
        femd_csr_s *a = check_csr(A);
        hipStream_t s = S(nullptr);
        const int n = a->pat->n;
        DevBuf bd = upload(b, size_t(n), s), xd = upload(x, size_t(n), s);
        refined_device(a, bd.as<double>(), xd.as<double>(), tol, inner_tol, max_outer, info, s);
        if (n) HIPC(hipMemcpyAsync(x, xd.p, 8 * size_t(n), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

// ---- timing --------------------------------------------------------------------------------
int femd_assembly_kernel_time(femd_space V, int quad_degree, int nterms, const int *a, const int *b,
                              const femd_coefficient *coeffs, int mode, int reps, double *ms, double *mfma_flops)
{
    return api([&] {
        femd_space_s *Vs = check_space(V);
        if (reps < 1 || mode < 0 || mode > 3 || !ms) fail(FEMD_EINVAL, "assembly_kernel_time: bad arguments");
        hipStream_t s = S(nullptr);
        ensure_pattern(Vs, s);
        Quad &Q = quad_of(Vs, quad_degree, s);
        Coeffs C = make_coeffs(Vs, Q, quad_degree, nterms, a, b, coeffs, s);
        const int n = Vs->V.nloc, nc = Vs->V.ncells;
        DevBuf E(8 * std::max<size_t>(size_t(nc) * n * n, 1));
        const MfmaPlan *mp = mode >= 1 ? mfma_plan(Vs, quad_degree, Q, nterms, a, b, coeffs, s, mode - 1) : nullptr;
        if (mode >= 1 && !mp) fail(FEMD_EINVAL, "assembly_kernel_time: the group is outside the matrix-core kernel's sizes");
        auto run = [&] {
            bool ok;
            if (mp && mode == 1)
                ok = launch_elem_matrix_mfma(Vs->V.nv, mfma_job(Vs, Q, *mp, E.as<double>()), C.T, mp->nt, mp->rlds, s);
            else if (mp)
                ok = launch_elem_matrix_mfma16(Vs->V.nv, mfma_job(Vs, Q, *mp, E.as<double>()), C.T, mp->nt, mode == 3, s);
            else
                ok = launch_elem_matrix(Vs->V.nv, n, nc, Q.nq, Vs->geom.as<double>(), Q.tab.as<double>(), Q.qw.as<double>(),
                                        Q.qxi.as<double>(), Q.qeta.as<double>(), C.T, E.as<double>(), s);
            if (!ok) fail(FEMD_EINVAL, "assembly_kernel_time: no element kernel for this group");
        };
        run();
        hipEvent_t e0, e1;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        HIPC(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) run();
        HIPC(hipEventRecord(e1, s));
        HIPC(hipEventSynchronize(e1));
        float f = 0.0f;
        HIPC(hipEventElapsedTime(&f, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = f / reps;
        // the matrix-core work issued: 16 cells x 16 entries x 4 k (2048 flops) per instruction
        if (mfma_flops) *mfma_flops = mp ? (mode == 1 ? 2048.0 : 8192.0) * double((nc + 15) / 16) * mp->ks * mp->ntile : 0.0;
    });
}

int femd_event_timer_start(femd_stream st)
{
    return api([&] {
        hipStream_t s = S(st);
        if (!t_ev0) { HIPC(hipEventCreate(&t_ev0)); HIPC(hipEventCreate(&t_ev1)); }
        HIPC(hipEventRecord(t_ev0, s));
    });
}

int femd_event_timer_stop(femd_stream st, double *ms)
{
    return api([&] {
        hipStream_t s = S(st);
        if (!t_ev0) fail(FEMD_EINVAL, "timer: not started");
        HIPC(hipEventRecord(t_ev1, s));
        HIPC(hipEventSynchronize(t_ev1));
        float f = 0.0f;
        HIPC(hipEventElapsedTime(&f, t_ev0, t_ev1));
        *ms = f;
    });
}

} // extern "C"
