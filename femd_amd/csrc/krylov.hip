// krylov.hip -- GMRES(restart) and LGMRES(inner, k_aug) with a left preconditioner
// (vendor/csnewton/csnewton.hpp:174-260, 298-474) behind the adapter of krylov/adapters.hpp:116-147,
// on the GPU.  Each Arnoldi step is enqueued without a host round trip: the modified Gram-Schmidt
// sweep, the new Hessenberg column, its Givens rotations and the residual estimate stay on the
// device (KState says when a cycle has converged, and every later kernel of the cycle returns at
// once); the host reads the state once per restart cycle, to choose the cycle's end.
//
// Below 4096 unknowns (FEMD_OMP_THRESHOLD, util/omp.hpp:25) the reference runs these solvers on one
// thread (krylov/adapters.hpp:50-68), so its dot products and norms are sums from left to right;
// there a whole Arnoldi step is one workgroup with w in LDS and every sum is formed in that order by
// one thread (the products in parallel, each rounded once as the reference's s += u[i] * v[i]), which
// makes iterates, iteration counts and residuals the reference's bit for bit.  Above, the reference's
// sums are OpenMP reductions whose bits depend on the thread count; here they are fixed-order tree
// reductions (deterministic run to run).
#include "cg_engine.hpp"

#include <cmath>

namespace femd_amd {

void spmv64(femd_csr_s *A, const double *x, double *y, const int *stop, hipStream_t s);
void precond_apply(femd_precond_s *M, const double *r, double *z, const int *stop, hipStream_t s);

namespace {

constexpr int KT = 1024;                  // threads of the one-workgroup kernels
constexpr int ORDERED_BELOW = 4096;       // FEMD_OMP_THRESHOLD
constexpr int TREE_THREADS = 256;
constexpr int TREE_BLOCKS = 512;

// The loop state of one cycle, on the device.
struct KState {
    double normb, tol, est;
    int conv;          // the cycle met the tolerance (or broke down) at step `last`
    int stop;          // every later kernel of the cycle returns at once
    int last;
    int breakdown;     // h_{j+1,j} == 0 at step `last`
};

// csnewton.hpp:126-148: the Givens rotation that zeroes dy, and its application
__device__ inline void givens_make(double dx, double dy, double &cs, double &sn)
{
    const double ax = fabs(dx), ay = fabs(dy);
    if (ay == 0.0) { cs = 1.0; sn = 0.0; }
    else if (ax == 0.0) { cs = 0.0; sn = 1.0; }
    else
    {
        const double tau = sqrt(ax * ax + ay * ay);
        cs = ax / tau;
        sn = (dx / ax) * (ay / tau);
    }
}
__device__ inline void givens_turn(double &dx, double &dy, double cs, double sn)
{
    const double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

// Column j of H: the earlier rotations, a new one, applied to H and to s; the residual estimate
// |s_{j+1}| / ||b|| against the tolerance (gmres: csnewton.hpp:220-235; lgmres: 363-372).
__device__ void rotate_column(int j, int ldh, double *H, double *cs, double *sn, double *s, double hn, bool lgmres, KState *st)
{
    double *col = H + size_t(j) * ldh;
    for (int k = 0; k < j; ++k) givens_turn(col[k], col[k + 1], cs[k], sn[k]);
    givens_make(col[j], col[j + 1], cs[j], sn[j]);
    givens_turn(col[j], col[j + 1], cs[j], sn[j]);
    givens_turn(s[j], s[j + 1], cs[j], sn[j]);
    const double est = fabs(s[j + 1]) / st->normb;
    if (est < st->tol || (lgmres && hn == 0.0))
    {
        st->conv = 1;
        st->stop = 1;
        st->last = j;
        st->est = est;
    }
}

// thread 0: 0 + p[0] + p[1] + ... from left to right
__device__ inline double serial_sum(const double *p, int n)
{
    double a = 0.0;
    for (int i = 0; i < n; ++i) a += p[i];
    return a;
}

// ||x|| with the serial sum (csnewton.hpp:100-107 on one thread; adapters.hpp:92-97), or a
// fixed-order tree sum of the block partials `parts` (ordered = 0; parts from k_tree_dot).
__global__ __launch_bounds__(KT) void k_norm_serial(int n, const double *__restrict__ x, double *out)
{
    __shared__ double p[ORDERED_BELOW];
    for (int i = threadIdx.x; i < n; i += KT) p[i] = x[i] * x[i];
    __syncthreads();
    if (threadIdx.x == 0) *out = sqrt(serial_sum(p, n));
}

// One Arnoldi step on one workgroup (n < 4096): w (the preconditioned product, left unchanged in
// HBM) against V[0..j] by modified Gram-Schmidt -- H[k][j] = w . V[k], then w += (-H[k][j]) V[k] --,
// h = ||w||, V[j+1] = w (1/h), and the column's rotations.  GMRES stops at h == 0 before any of the
// last three (its solution update then uses the column as it stands, csnewton.hpp:211-214).
template <bool LGMRES>
__global__ __launch_bounds__(KT) void k_arnoldi_serial(int n, int j, int ldh, double *V, const double *__restrict__ win,
                                                       double *H, double *cs, double *sn, double *s, KState *st)
{
    __shared__ double w[ORDERED_BELOW], p[ORDERED_BELOW];
    __shared__ double h;
    if (st->stop) return;
    const int t = threadIdx.x;
    for (int i = t; i < n; i += KT) w[i] = win[i];
    for (int k = 0; k <= j; ++k)
    {
        const double *vk = V + size_t(k) * n;
        for (int i = t; i < n; i += KT) p[i] = w[i] * vk[i];
        __syncthreads();
        if (t == 0)
        {
            h = serial_sum(p, n);
            H[k + size_t(j) * ldh] = h;
        }
        __syncthreads();
        const double a = -h;
        for (int i = t; i < n; i += KT) w[i] = w[i] + a * vk[i];
        __syncthreads();
    }
    for (int i = t; i < n; i += KT) p[i] = w[i] * w[i];
    __syncthreads();
    if (t == 0)
    {
        h = sqrt(serial_sum(p, n));
        H[(j + 1) + size_t(j) * ldh] = h;
    }
    __syncthreads();
    const double hn = h;
    if (!LGMRES && hn == 0.0)
    {
        if (t == 0) { st->conv = 1; st->stop = 1; st->last = j; st->breakdown = 1; st->est = 0.0; }
        return;
    }
    if (hn > 0.0)
    {
        const double a = 1.0 / hn;
        double *vn = V + size_t(j + 1) * n;
        for (int i = t; i < n; i += KT) vn[i] = w[i] * a;
    }
    if (t == 0) rotate_column(j, ldh, H, cs, sn, s, hn, LGMRES, st);
}

// ---- the tree path (n >= 4096): the same step as a sequence of kernels
__device__ double block_reduce(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    v = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < int(blockDim.x >> 6); ++k) v += sh[k];
    return v;
}

__global__ __launch_bounds__(TREE_THREADS) void k_tree_dot(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                           double *parts, const KState *st)
{
    __shared__ double sh[TREE_THREADS / 64];
    if (st && st->stop) return;
    double a = 0.0;
    for (int i = blockIdx.x * TREE_THREADS + threadIdx.x; i < n; i += TREE_BLOCKS * TREE_THREADS) a += x[i] * y[i];
    a = block_reduce(a, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = a;
}

// out = the partials' sum (root: its square root), in a fixed order
__global__ __launch_bounds__(TREE_THREADS) void k_tree_sum(const double *parts, double *out, int root, const KState *st)
{
    __shared__ double sh[TREE_THREADS / 64];
    if (st && st->stop) return;
    double a = 0.0;
    for (int i = threadIdx.x; i < TREE_BLOCKS; i += TREE_THREADS) a += parts[i];
    a = block_reduce(a, sh);
    if (threadIdx.x == 0) *out = root ? sqrt(a) : a;
}

// w += (-h) v
__global__ void k_tree_axpy(int n, const double *h, const double *__restrict__ v, double *w, const KState *st)
{
    if (st->stop) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = w[i] + (-*h) * v[i];
}

// the step's end: GMRES breakdown, V[j+1] = w (1/h), the rotations (one block; thread 0 rotates)
__global__ void k_tree_next(int n, int j, int ldh, double *V, const double *__restrict__ w, double *H, double *cs, double *sn,
                            double *s, int lgmres, KState *st)
{
    if (st->stop) return;
    const double hn = H[(j + 1) + size_t(j) * ldh];
    if (!lgmres && hn == 0.0)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) { st->conv = 1; st->stop = 1; st->last = j; st->breakdown = 1; st->est = 0.0; }
        return;
    }
    if (hn > 0.0)
    {
        const double a = 1.0 / hn;
        double *vn = V + size_t(j + 1) * n;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) vn[i] = w[i] * a;
    }
}
__global__ void k_tree_rotate(int j, int ldh, double *H, double *cs, double *sn, double *s, int lgmres, KState *st)
{
    if (st->stop) return;
    rotate_column(j, ldh, H, cs, sn, s, H[(j + 1) + size_t(j) * ldh], lgmres != 0, st);
}

// ---- shared by both paths
__global__ void k_copy_unless(int n, const double *__restrict__ a, double *b, const KState *st)
{
    if (st && st->stop) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// V0 = q (1 / beta), the cycle's first basis vector and s = (beta, 0, ...) (csnewton.hpp:196-200, 339-346)
__global__ void k_cycle_start(int n, const double *__restrict__ q, const double *beta, double *V0, double *s, int ns, double *H,
                              int nh)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double a = 1.0 / *beta;
    if (i < n) V0[i] = q[i] * a;
    if (i < ns) s[i] = i == 0 ? *beta : 0.0;
    for (int k = i; k < nh; k += gridDim.x * blockDim.x) H[k] = 0.0;
}

// y = the back-substituted s (update_solution, csnewton.hpp:150-155; lgmres 399-405), one thread
__global__ void k_back_solve(int k, int ldh, const double *H, const double *s, double *y)
{
    for (int i = 0; i <= k; ++i) y[i] = s[i];
    for (int i = k; i >= 0; --i)
    {
        y[i] /= H[i + size_t(i) * ldh];
        for (int jj = 0; jj < i; ++jj) y[jj] -= H[jj + size_t(i) * ldh] * y[i];
    }
}

struct Basis { const double *v[64]; };
// x[i] += y_j B_j[i] for j in order (and z[i] += the same terms, when z is given: lgmres, csnewton.hpp:410-418)
__global__ void k_combine(int n, int k, const double *y, Basis B, double *x, double *z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xi = x[i], zi = z ? z[i] : 0.0;
    for (int j = 0; j < k; ++j)
    {
        const double c = y[j] * B.v[j][i];
        xi += c;
        zi += c;
    }
    x[i] = xi;
    if (z) z[i] = zi;
}

__global__ void k_residual(int n, const double *__restrict__ b, const double *__restrict__ Ax, double *r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - Ax[i];
}

int grid(int n, int t = 256) { return (n + t - 1) / t > 0 ? (n + t - 1) / t : 1; }

// One solve's device work space and the operations of both paths.
struct Run {
    femd_csr_s *A;
    femd_precond_s *M;
    hipStream_t s;
    int n;
    bool serial;
    DevBuf parts, scal, stb;
    KState *st;
    Run(femd_csr_s *A_, femd_precond_s *M_, hipStream_t s_)
        : A(A_), M(M_), s(s_), n(A_->pat->n), serial(A_->pat->n < ORDERED_BELOW), parts(8 * TREE_BLOCKS), scal(64),
          stb(sizeof(KState))
    {
        st = stb.as<KState>();
    }
    void matvec(const double *x, double *y, bool guarded) { spmv64(A, x, y, guarded ? &st->stop : nullptr, s); }
    void precon(const double *r, double *z, bool guarded)
    {
        if (M) precond_apply(M, r, z, guarded ? &st->stop : nullptr, s);
        else hipLaunchKernelGGL(k_copy_unless, dim3(grid(n)), dim3(256), 0, s, n, r, z, guarded ? st : nullptr);
    }
    // ||x|| into the device scalar out
    void norm_into(const double *x, double *out)
    {
        if (serial) hipLaunchKernelGGL(k_norm_serial, dim3(1), dim3(KT), 0, s, n, x, out);
        else
        {
            hipLaunchKernelGGL(k_tree_dot, dim3(TREE_BLOCKS), dim3(TREE_THREADS), 0, s, n, x, x, parts.as<double>(), nullptr);
            hipLaunchKernelGGL(k_tree_sum, dim3(1), dim3(TREE_THREADS), 0, s, parts.as<double>(), out, 1, nullptr);
        }
    }
    double norm(const double *x)
    {
        norm_into(x, scal.as<double>());
        return read_scalar(scal.as<double>(), s);
    }
    void reset(double normb, double tol)
    {
        KState h{};
        h.normb = normb;
        h.tol = tol;
        h.last = -1;
        HIPC(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    }
    KState state()
    {
        KState h{};
        HIPC(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        return h;
    }
    // Arnoldi step j on input w (the preconditioned product, or a stored M^{-1} A z)
    void arnoldi(int j, int ldh, double *V, const double *w, double *work, double *H, double *cs, double *sn, double *sv,
                 bool lgmres)
    {
        if (serial)
        {
            if (lgmres) hipLaunchKernelGGL(k_arnoldi_serial<true>, dim3(1), dim3(KT), 0, s, n, j, ldh, V, w, H, cs, sn, sv, st);
            else hipLaunchKernelGGL(k_arnoldi_serial<false>, dim3(1), dim3(KT), 0, s, n, j, ldh, V, w, H, cs, sn, sv, st);
            return;
        }
        hipLaunchKernelGGL(k_copy_unless, dim3(grid(n)), dim3(256), 0, s, n, w, work, st);
        double *col = H + size_t(j) * ldh;
        for (int k = 0; k <= j; ++k)
        {
            const double *vk = V + size_t(k) * n;
            hipLaunchKernelGGL(k_tree_dot, dim3(TREE_BLOCKS), dim3(TREE_THREADS), 0, s, n, work, vk, parts.as<double>(), st);
            hipLaunchKernelGGL(k_tree_sum, dim3(1), dim3(TREE_THREADS), 0, s, parts.as<double>(), col + k, 0, st);
            hipLaunchKernelGGL(k_tree_axpy, dim3(grid(n)), dim3(256), 0, s, n, col + k, vk, work, st);
        }
        hipLaunchKernelGGL(k_tree_dot, dim3(TREE_BLOCKS), dim3(TREE_THREADS), 0, s, n, work, work, parts.as<double>(), st);
        hipLaunchKernelGGL(k_tree_sum, dim3(1), dim3(TREE_THREADS), 0, s, parts.as<double>(), col + j + 1, 1, st);
        hipLaunchKernelGGL(k_tree_next, dim3(std::min(grid(n), 1024)), dim3(256), 0, s, n, j, ldh, V, work, H, cs, sn, sv,
                           int(lgmres), st);
        hipLaunchKernelGGL(k_tree_rotate, dim3(1), dim3(1), 0, s, j, ldh, H, cs, sn, sv, int(lgmres), st);
    }
    // r = b - A x, q = M^{-1} r, ||q|| into out
    void residual(const double *x, const double *b, double *Ax, double *r, double *q, double *out)
    {
        matvec(x, Ax, false);
        hipLaunchKernelGGL(k_residual, dim3(grid(n)), dim3(256), 0, s, n, b, Ax, r);
        precon(r, q, false);
        norm_into(q, out);
    }
    void combine(int k, const double *y, const std::vector<const double *> &B, double *x, double *z)
    {
        for (int j0 = 0; j0 < k; j0 += 64)
        {
            Basis P{};
            const int kk = std::min(64, k - j0);
            for (int j = 0; j < kk; ++j) P.v[j] = B[j0 + j];
            hipLaunchKernelGGL(k_combine, dim3(grid(n)), dim3(256), 0, s, n, kk, y + j0, P, x, z);
        }
    }
};

struct Outcome { bool converged; int iterations; double residual; };

// csnewton.hpp:174-260
Outcome gmres_run(Run &R, double *x, const double *b, int restart, int max_iter, double tol)
{
    const int n = R.n, ldh = restart + 1;
    DevBuf Vb(8 * size_t(n) * (restart + 1)), w(8 * size_t(n)), r(8 * size_t(n)), z(8 * size_t(n)), work(8 * size_t(n));
    DevBuf H(8 * size_t(ldh) * restart), cs(8 * size_t(restart)), sn(8 * size_t(restart)), sv(8 * size_t(restart + 1)),
        y(8 * size_t(restart + 1)), beta(8);
    double *V = Vb.as<double>();
    std::vector<const double *> basis(restart + 1);
    for (int i = 0; i <= restart; ++i) basis[i] = V + size_t(i) * n;
    R.residual(x, b, w.as<double>(), r.as<double>(), z.as<double>(), beta.as<double>());
    double normb = R.norm(b);
    if (normb == 0.0) normb = 1.0;
    double bt = read_scalar(beta.as<double>(), R.s);
    if (bt / normb < tol) return {true, 0, bt / normb};
    int total = 0;
    while (total < max_iter)
    {
        hipLaunchKernelGGL(k_cycle_start, dim3(grid(std::max(n, restart + 1))), dim3(256), 0, R.s, n, z.as<double>(),
                           beta.as<double>(), V, sv.as<double>(), restart + 1, H.as<double>(), ldh * restart);
        R.reset(normb, tol);
        const int steps = std::min(restart, max_iter - total);
        for (int i = 0; i < steps; ++i)
        {
            R.matvec(basis[i], r.as<double>(), true);
            R.precon(r.as<double>(), w.as<double>(), true);
            R.arnoldi(i, ldh, V, w.as<double>(), work.as<double>(), H.as<double>(), cs.as<double>(), sn.as<double>(),
                      sv.as<double>(), false);
        }
        const KState st = R.state();
        if (st.conv)
        {
            hipLaunchKernelGGL(k_back_solve, dim3(1), dim3(1), 0, R.s, st.last, ldh, H.as<double>(), sv.as<double>(), y.as<double>());
            R.combine(st.last + 1, y.as<double>(), basis, x, nullptr);
            HIPC(hipStreamSynchronize(R.s));
            return {true, total + st.last + 1, st.breakdown ? 0.0 : st.est};
        }
        total += steps;
        hipLaunchKernelGGL(k_back_solve, dim3(1), dim3(1), 0, R.s, steps - 1, ldh, H.as<double>(), sv.as<double>(), y.as<double>());
        R.combine(steps, y.as<double>(), basis, x, nullptr);
        R.residual(x, b, w.as<double>(), r.as<double>(), z.as<double>(), beta.as<double>());
        bt = read_scalar(beta.as<double>(), R.s);
        if (bt / normb < tol) return {true, total, bt / normb};
    }
    return {false, total, bt / normb};
}

// csnewton.hpp:298-474
Outcome lgmres_run(Run &R, double *x, const double *b, int inner, int k_aug, int max_iter, double tol)
{
    const int n = R.n;
    if (k_aug < 0) fail(FEMD_EINVAL, "lgmres: k_aug must be non-negative");
    const int mmax = inner + k_aug, slots = std::max(k_aug, 1);
    DevBuf Vb(8 * size_t(n) * (mmax + 1)), Zb(8 * size_t(n) * slots), MZb(8 * size_t(n) * slots);
    DevBuf Ax(8 * size_t(n)), r(8 * size_t(n)), q(8 * size_t(n)), w(8 * size_t(n)), zn(8 * size_t(n)), work(8 * size_t(n));
    DevBuf H(8 * size_t(mmax + 1) * mmax), cs(8 * size_t(mmax)), sn(8 * size_t(mmax)), sv(8 * size_t(mmax + 1)),
        y(8 * size_t(mmax + 1)), beta(8);
    double *V = Vb.as<double>();
    std::vector<double *> zslot(slots), mslot(slots), Zs, MZs;      // Zs[0] is the newest correction
    for (int i = 0; i < slots; ++i)
    {
        zslot[i] = Zb.as<double>() + size_t(i) * n;
        mslot[i] = MZb.as<double>() + size_t(i) * n;
    }
    double normb = R.norm(b);
    if (normb == 0.0) normb = 1.0;
    int total = 0;
    auto final_residual = [&]() {
        R.residual(x, b, Ax.as<double>(), r.as<double>(), q.as<double>(), beta.as<double>());
        return read_scalar(beta.as<double>(), R.s) / normb;
    };
    while (total < max_iter)
    {
        const int k = static_cast<int>(Zs.size()), m = inner + k, ldh = m + 1;
        R.residual(x, b, Ax.as<double>(), r.as<double>(), q.as<double>(), beta.as<double>());
        const double bt = read_scalar(beta.as<double>(), R.s);
        if (bt / normb < tol) return {true, total, bt / normb};
        hipLaunchKernelGGL(k_cycle_start, dim3(grid(std::max(n, m + 1))), dim3(256), 0, R.s, n, q.as<double>(), beta.as<double>(),
                           V, sv.as<double>(), m + 1, H.as<double>(), ldh * m);
        R.reset(normb, tol);
        // phase 1: inner Arnoldi steps; phase 2: the stored corrections (their M^{-1} A z kept)
        const int inner_lim = std::min(inner, max_iter - total);
        for (int j = 0; j < inner_lim; ++j)
        {
            R.matvec(V + size_t(j) * n, r.as<double>(), true);
            R.precon(r.as<double>(), w.as<double>(), true);
            R.arnoldi(j, ldh, V, w.as<double>(), work.as<double>(), H.as<double>(), cs.as<double>(), sn.as<double>(),
                      sv.as<double>(), true);
        }
        const int aug_lim = std::min(k, max_iter - (total + inner_lim));
        for (int a = 0; a < aug_lim; ++a)
            R.arnoldi(inner + a, ldh, V, MZs[a], work.as<double>(), H.as<double>(), cs.as<double>(), sn.as<double>(),
                      sv.as<double>(), true);
        const KState st = R.state();
        int final_j;
        if (st.conv)
        {
            // the steps before it and its own extra count (csnewton.hpp:374, 395); phase 2 runs only
            // after all `inner` steps of phase 1, so step j is the cycle's (j + 1)-th either way
            final_j = st.last;
            total += st.last + 1;
        }
        else
        {
            total += inner_lim + aug_lim;
            final_j = inner_lim + aug_lim - 1;
        }
        hipLaunchKernelGGL(k_back_solve, dim3(1), dim3(1), 0, R.s, final_j, ldh, H.as<double>(), sv.as<double>(), y.as<double>());
        std::vector<const double *> basis(final_j + 1);
        for (int j = 0; j <= final_j; ++j) basis[j] = j < inner ? V + size_t(j) * n : Zs[j - inner];
        HIPC(hipMemsetAsync(zn.p, 0, 8 * size_t(n), R.s));
        R.combine(final_j + 1, y.as<double>(), basis, x, zn.as<double>());
        if (k_aug > 0)
        {
            // the newest correction takes a free slot, or the oldest one's
            double *zs, *ms;
            if (static_cast<int>(Zs.size()) < k_aug) { zs = zslot[Zs.size()]; ms = mslot[Zs.size()]; }
            else { zs = Zs.back(); ms = MZs.back(); Zs.pop_back(); MZs.pop_back(); }
            hipLaunchKernelGGL(k_copy_unless, dim3(grid(n)), dim3(256), 0, R.s, n, zn.as<double>(), zs, nullptr);
            R.matvec(zs, Ax.as<double>(), false);
            R.precon(Ax.as<double>(), ms, false);
            Zs.insert(Zs.begin(), zs);
            MZs.insert(MZs.begin(), ms);
        }
        HIPC(hipStreamSynchronize(R.s));
        if (st.conv) return {true, total, final_residual()};
    }
    return {false, total, final_residual()};
}

} // namespace

// krylov/adapters.hpp:116-147 around the solvers: tol rescaled to tol ||M^{-1} b|| / ||b||, the true
// and the preconditioned residuals recomputed at the end, and the verdict.
void krylov_device(femd_csr_s *A, femd_precond_s *M, const double *b, double *x, int lgmres, int restart, int max_iter,
                   double tol, int k_aug, femd_solve_info *info, hipStream_t s)
{
    const int n = A->pat->n;
    std::memset(info, 0, sizeof(*info));
    Run R(A, M, s);
    const double nb = R.norm(b);
    if (nb == 0.0)
    {
        launch_fill(n, 0.0, x, s);
        HIPC(hipStreamSynchronize(s));
        info->converged = 1;
        return;
    }
    DevBuf Mb(8 * size_t(std::max(n, 1)));
    R.precon(b, Mb.as<double>(), false);
    const double nMb = R.norm(Mb.as<double>());
    const double scaled = nMb > 0.0 ? tol * nMb / nb : tol;
    const Outcome g = lgmres ? lgmres_run(R, x, b, restart, k_aug, max_iter, scaled) : gmres_run(R, x, b, restart, max_iter, scaled);
    DevBuf Ax(8 * size_t(n)), r(8 * size_t(n)), Mr(8 * size_t(n));
    R.matvec(x, Ax.as<double>(), false);
    hipLaunchKernelGGL(k_residual, dim3(grid(n)), dim3(256), 0, s, n, b, Ax.as<double>(), r.as<double>());
    R.precon(r.as<double>(), Mr.as<double>(), false);
    info->iterations = g.iterations;
    info->residual = R.norm(r.as<double>()) / nb;
    info->precond_residual = nMb > 0.0 ? R.norm(Mr.as<double>()) / nMb : 0.0;
    info->converged = info->precond_residual <= tol * (1.0 + 1e-8) || (g.converged && info->precond_residual <= 10.0 * tol);
}

} // namespace femd_amd
