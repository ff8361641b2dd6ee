// ref_driver.cpp -- a C ABI over the FEMd reference headers, for the test suite only.
//
// This file is test infrastructure.  It is compiled (oracle/Makefile, target `ref`) against
// the reference's own header-only sources where they lie, and the shared library goes to
// oracle/_ref/, which is never committed.  Nothing in the product links or loads it; the
// tests and the CPU-baseline leg of bench.py use it as the checker and as the CPU timing of
// the reference's own code.
//
// Every entry point returns 0 on success and -1 on a C++ exception, whose text is kept for
// ref_last_error().  Handles are opaque pointers owned by the caller (ref_*_free).
//
// What each entry point drives in the reference:
//   ref_mesh_rectangle      mesh/structured.hpp:338-351 (rectangle_mesh)
//   ref_mesh_from_arrays    mesh/structured.hpp:99-221  (mesh_from_arrays)
//   ref_space_create        fe/lagrange_space_2d.hpp:32-67, fe/space_2d.hpp:372-526
//   ref_quad_space_create   mesh/quad_mesh.hpp:50-103, fe/lagrange_space_quad.hpp:112-210
//   ref_quad_quadrature, ref_quad_basis_eval   fe/cache_2d.hpp:75-84, lagrange_space_quad.hpp:28-110
//   ref_pattern             forms/assembly_2d.hpp:36-41, sparse/csr.hpp:84-125
//   ref_assemble_matrix     fe/cache_2d.hpp:67-121, forms/assembly_2d.hpp:136-164,219-305
//   ref_assemble_vector     forms/assembly_2d.hpp:173-211
//   ref_integrand_eval      forms/integrand.hpp:85-109,300-342 (ref_integrand_eval_in: with inputs)
//   ref_eval_field          fe/cache_2d.hpp:245-288 (Cache2D::eval_multi)
//   ref_spmv / ref_ddot     sparse/csr.hpp:203-215, util/omp.hpp:176-192
//   ref_cg                  krylov/cg.hpp:245-350 as python/src/femd_bindings.cpp:799-833 calls it
//   ref_krylov              krylov/adapters.hpp:116-147 (csnewton gmres / lgmres)
//   ref_rcm, ref_precond_apply  sparse/amd.hpp:166-234, sparse/precond.hpp:82-205
//   ref_asymmetry           sparse/csr.hpp:631-644
#include "femd/femd.hpp"
#include "femd/fe/lagrange_space_quad.hpp"
#include "femd/krylov/adapters.hpp"
#include "femd/mesh/quad_mesh.hpp"

#include <cstring>
#include <omp.h>
#include <memory>
#include <string>
#include <vector>

using namespace femd;

namespace {
thread_local std::string g_err;

template <class F> int guard(F &&f)
{
    try { f(); return 0; }
    catch (const std::exception &e) { g_err = e.what(); return -1; }
    catch (...) { g_err = "unknown exception"; return -1; }
}

struct Space {
    std::shared_ptr<Mesh2D> mesh;
    std::shared_ptr<QuadMesh> qmesh;
    std::unique_ptr<Space2D> V;
};

CSRMatrix make_csr(int n, int m, const int *ip, const int *ix, const double *v)
{
    auto P = std::make_shared<SparsityPattern>();
    P->nrows = n; P->ncols = m;
    P->indptr.assign(ip, ip + n + 1);
    P->indices.assign(ix, ix + ip[n]);
    CSRMatrix K(P);
    K.data.assign(v, v + ip[n]);
    return K;
}
} // namespace

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

/// The OpenMP team size of the reference's threaded loops (its vendored GMRES dots are
/// OpenMP reductions whose bits depend on it; everything else is thread-count independent).
void ref_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void *ref_mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, const char *diag)
{
    Mesh2D *out = nullptr;
    guard([&] { out = new Mesh2D(rectangle_mesh(x0, y0, x1, y1, nx, ny, diagonal_from_name(diag))); });
    return out;
}

void *ref_mesh_from_arrays(int np, const double *pts, int nt, const int *tris, int ns, const int *segs, const int *smarks)
{
    Mesh2D *out = nullptr;
    guard([&] {
        std::vector<Point> P(static_cast<std::size_t>(np));
        for (int i = 0; i < np; ++i) P[i] = Point{pts[2 * i], pts[2 * i + 1]};
        std::vector<std::array<int, 3>> T(static_cast<std::size_t>(nt));
        for (int t = 0; t < nt; ++t) T[t] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        std::vector<std::array<int, 2>> S(static_cast<std::size_t>(ns));
        std::vector<int> M(static_cast<std::size_t>(ns));
        for (int s = 0; s < ns; ++s) { S[s] = {segs[2 * s], segs[2 * s + 1]}; M[s] = smarks ? smarks[s] : 1; }
        out = new Mesh2D(mesh_from_arrays(P, T, S, smarks ? M : std::vector<int>{}));
    });
    return out;
}

void ref_mesh_free(void *m) { delete static_cast<Mesh2D *>(m); }

int ref_mesh_sizes(void *mh, int *out3)
{
    return guard([&] {
        auto *m = static_cast<Mesh2D *>(mh);
        out3[0] = m->npoints(); out3[1] = m->ntriangles(); out3[2] = m->nsegments();
    });
}

int ref_mesh_arrays(void *mh, double *pts, int *tris, int *nbr, int *segs, int *smarks)
{
    return guard([&] {
        auto *m = static_cast<Mesh2D *>(mh);
        for (int i = 0; i < m->npoints(); ++i) { pts[2 * i] = m->point(i).x; pts[2 * i + 1] = m->point(i).y; }
        for (int t = 0; t < m->ntriangles(); ++t)
            for (int i = 0; i < 3; ++i) { tris[3 * t + i] = m->triangle(t)[i]; nbr[3 * t + i] = m->neighbour(t, i); }
        for (int s = 0; s < m->nsegments(); ++s)
        {
            segs[2 * s] = m->segments()[s][0]; segs[2 * s + 1] = m->segments()[s][1];
            smarks[s] = m->segment_markers()[s];
        }
    });
}

void *ref_space_create(void *mh, int k, int ndir, const int *dir, int all_exterior, int equispaced)
{
    Space *out = nullptr;
    guard([&] {
        auto S = std::make_unique<Space>();
        S->mesh = std::make_shared<Mesh2D>(*static_cast<Mesh2D *>(mh));
        std::vector<int> d(dir, dir + ndir);
        S->V = std::make_unique<LagrangeSpace2D>(*S->mesh, k, d, all_exterior != 0, equispaced != 0);
        out = S.release();
    });
    return out;
}

/// A QuadMesh from arrays (quads counter-clockwise or clockwise; the listed segments with their
/// markers, other boundary edges marker 1) and a Q_k space on it.
void *ref_quad_space_create(int np, const double *pts, int nq, const int *quads, int ns, const int *segs, const int *smarks,
                            int k, int ndir, const int *dir, int all_exterior, int equispaced)
{
    Space *out = nullptr;
    guard([&] {
        auto S = std::make_unique<Space>();
        std::vector<Point> P(static_cast<std::size_t>(np));
        for (int i = 0; i < np; ++i) P[i] = Point{pts[2 * i], pts[2 * i + 1]};
        std::vector<std::array<int, 4>> Q(static_cast<std::size_t>(nq));
        for (int q = 0; q < nq; ++q) Q[q] = {quads[4 * q], quads[4 * q + 1], quads[4 * q + 2], quads[4 * q + 3]};
        std::vector<std::array<int, 2>> G(static_cast<std::size_t>(ns));
        for (int e = 0; e < ns; ++e) G[e] = {segs[2 * e], segs[2 * e + 1]};
        std::vector<int> M = smarks ? std::vector<int>(smarks, smarks + ns) : std::vector<int>{};
        S->qmesh = std::make_shared<QuadMesh>(std::move(P), std::move(Q), std::move(G), std::move(M), std::vector<int>{}, 1);
        S->V = std::make_unique<LagrangeSpaceQ>(*S->qmesh, k, std::vector<int>(dir, dir + ndir), all_exterior != 0, equispaced != 0);
        out = S.release();
    });
    return out;
}

/// The quads of a space's mesh after orientation (nq x 4).
int ref_quad_mesh_quads(void *sh, int *quads)
{
    return guard([&] {
        const QuadMesh &M = *static_cast<Space *>(sh)->qmesh;
        for (int q = 0; q < M.nquads(); ++q)
            for (int i = 0; i < 4; ++i) quads[4 * q + i] = M.quad(q)[i];
    });
}

/// The quadrilateral rule Cache2D uses for `degree` (up to 64 points).
int ref_quad_quadrature(int degree, int *nq, double *xi, double *eta, double *w)
{
    return guard([&] {
        GaussLegendre g(std::max(1, (degree + 3) / 2));
        int k = 0;
        for (int j = 0; j < g.size(); ++j)
            for (int i = 0; i < g.size(); ++i, ++k)
            {
                xi[k] = 0.5 * (g.node(i) + 1.0); eta[k] = 0.5 * (g.node(j) + 1.0);
                w[k] = 0.25 * g.weight(i) * g.weight(j);
            }
        *nq = k;
    });
}

/// Values, d/dxi, d/deta of the Q_k basis at (xi, eta): 3 (k+1)^2 numbers.
int ref_quad_basis_eval(int k, int equispaced, double xi, double eta, double *out)
{
    return guard([&] {
        ReferenceQuad R(k, equispaced != 0);
        R.eval(xi, eta, 1, out);
    });
}

void ref_space_free(void *s) { delete static_cast<Space *>(s); }

/// out: ncells, nloc, raw_dim, dim, nedges, npoints
int ref_space_sizes(void *sh, int *out)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        out[0] = V.ncells(); out[1] = V.nloc(); out[2] = V.raw_dim(); out[3] = V.dim(); out[4] = V.nedges(); out[5] = V.npoints();
    });
}

int ref_space_numbering(void *sh, int *cell_dof, int *cell_raw, int *ad2raw, double *raw_xy)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        std::copy(V.cell_dofs().begin(), V.cell_dofs().end(), cell_dof);
        std::copy(V.cell_raw_dofs().begin(), V.cell_raw_dofs().end(), cell_raw);
        for (int a = 0; a < V.dim(); ++a) ad2raw[a] = V.adapted_to_raw(a);
        std::copy(V.raw_coordinates().begin(), V.raw_coordinates().end(), raw_xy);
    });
}

long ref_pattern_nnz(void *sh)
{
    long n = -1;
    guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        n = cell_pattern(V, V)->nnz();
    });
    return n;
}

int ref_pattern(void *sh, int *indptr, int *indices)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        auto P = cell_pattern(V, V);
        std::copy(P->indptr.begin(), P->indptr.end(), indptr);
        std::copy(P->indices.begin(), P->indices.end(), indices);
    });
}

/// The triangle rule of `degree`: nq on return (arrays hold up to 64 points).
int ref_quadrature(int degree, int *nq, double *xi, double *eta, double *w)
{
    return guard([&] {
        TriangleQuadrature R(degree);
        *nq = R.size();
        for (int q = 0; q < R.size(); ++q) { xi[q] = R.xi(q); eta[q] = R.eta(q); w[q] = R.weight(q); }
    });
}

/// Values, d/dxi, d/deta of the P_k basis at (xi, eta): 3 * nloc numbers.
int ref_basis_eval(int k, int equispaced, double xi, double eta, double *out)
{
    return guard([&] {
        ReferenceTriangle T(k, equispaced != 0);
        T.eval(xi, eta, 1, out);
    });
}

/// The local nodes of ReferenceTriangle (quad = 0) or ReferenceQuad (quad = 1).
int ref_nodes(int k, int equispaced, int quad, double *xi, double *eta)
{
    return guard([&] {
        if (quad)
        {
            ReferenceQuad R(k, equispaced != 0);
            for (int l = 0; l < R.size(); ++l) { xi[l] = R.node_xi(l); eta[l] = R.node_eta(l); }
        }
        else
        {
            ReferenceTriangle R(k, equispaced != 0);
            for (int l = 0; l < R.size(); ++l) { xi[l] = R.node_xi(l); eta[l] = R.node_eta(l); }
        }
    });
}

/// detail::gauss_lobatto_nodes (fe/dg_space.hpp:60-80).
int ref_lobatto_points(int p, double *out)
{
    return guard([&] {
        const std::vector<double> x = detail::gauss_lobatto_nodes(p);
        std::copy(x.begin(), x.end(), out);
    });
}

/// Quadrature points of the cell cache: x, y, w (ncells * nq each).
int ref_cache_points(void *sh, int degree, double *x, double *y, double *w)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        Cache2D Q(V, degree);
        std::copy(Q.x().begin(), Q.x().end(), x);
        std::copy(Q.y().begin(), Q.y().end(), y);
        std::copy(Q.weights().begin(), Q.weights().end(), w);
    });
}

/// K (zeroed here) = sum_t integral c_t (D^{a_t} phi_i)(D^{b_t} phi_j) on the space's own pattern.
/// coeffs: nterms arrays of ncells * nq values, laid out one after another.
int ref_assemble_matrix(void *sh, int degree, int nterms, const int *a, const int *b, const double *coeffs, double *data)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        Cache2D Q(V, degree);
        const std::size_t np = static_cast<std::size_t>(Q.nent()) * Q.nq();
        std::vector<std::vector<double>> c(static_cast<std::size_t>(nterms));
        for (int t = 0; t < nterms; ++t) c[t].assign(coeffs + t * np, coeffs + (t + 1) * np);
        CSRMatrix K(cell_pattern(V, V));
        std::vector<int> A(a, a + nterms), B(b, b + nterms);
        std::vector<int> map = scatter_map_2d(Q, Q, *K.pattern, 0, 0);
        assemble_matrix_2d(Q, Q, A, B, c, K, 0, 0, map.data());
        std::copy(K.data.begin(), K.data.end(), data);
    });
}

int ref_assemble_vector(void *sh, int degree, int nterms, const int *a, const double *coeffs, double *f)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        Cache2D Q(V, degree);
        const std::size_t np = static_cast<std::size_t>(Q.nent()) * Q.nq();
        std::vector<std::vector<double>> c(static_cast<std::size_t>(nterms));
        for (int t = 0; t < nterms; ++t) c[t].assign(coeffs + t * np, coeffs + (t + 1) * np);
        std::vector<double> block(static_cast<std::size_t>(V.dim()), 0.0);
        assemble_vector_2d(Q, std::vector<int>(a, a + nterms), c, block);
        for (int i = 0; i < V.dim(); ++i) f[i] = 0.0 + block[i];
    });
}

/// The reference's ds point set: Cache2D(V, degree, markers, all) (fe/cache_2d.hpp:124-183).
int ref_facet_sizes(void *sh, int degree, int nmark, const int *marks, int all, int *out)
{
    return guard([&] {
        Cache2D Q(*static_cast<Space *>(sh)->V, degree, std::vector<int>(marks, marks + nmark), all != 0);
        out[0] = Q.nent(); out[1] = Q.nq();
    });
}

int ref_facet_points(void *sh, int degree, int nmark, const int *marks, int all, double *x, double *y, double *w, double *nx,
                     double *ny)
{
    return guard([&] {
        Cache2D Q(*static_cast<Space *>(sh)->V, degree, std::vector<int>(marks, marks + nmark), all != 0);
        std::copy(Q.x().begin(), Q.x().end(), x);
        std::copy(Q.y().begin(), Q.y().end(), y);
        std::copy(Q.weights().begin(), Q.weights().end(), w);
        std::copy(Q.normal_x().begin(), Q.normal_x().end(), nx);
        std::copy(Q.normal_y().begin(), Q.normal_y().end(), ny);
    });
}

/// A form's dx / ds groups added into one matrix (vector) as FormAssembler2D does
/// (forms/assembler_2d.hpp:166-188, 249-266), each group through the reference's own cache,
/// scatter map and assembly routine.
int ref_assemble_matrix_groups(void *sh, int ngroups, const int *kind, const int *degree, const int *nmark, const int *marks,
                               const int *all, const int *nterms, const int *a, const int *b, const double *coeffs, double *data)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        CSRMatrix K(cell_pattern(V, V));
        std::size_t mo = 0, to = 0, co = 0;
        for (int g = 0; g < ngroups; ++g)
        {
            std::unique_ptr<Cache2D> Q = kind[g] == 0 ? std::make_unique<Cache2D>(V, degree[g])
                                                      : std::make_unique<Cache2D>(V, degree[g], std::vector<int>(marks + mo, marks + mo + nmark[g]), all[g] != 0);
            const std::size_t np = static_cast<std::size_t>(Q->nent()) * Q->nq();
            std::vector<std::vector<double>> c(static_cast<std::size_t>(nterms[g]));
            for (int t = 0; t < nterms[g]; ++t) c[t].assign(coeffs + co + t * np, coeffs + co + (t + 1) * np);
            std::vector<int> A(a + to, a + to + nterms[g]), B(b + to, b + to + nterms[g]);
            std::vector<int> map = scatter_map_2d(*Q, *Q, *K.pattern, 0, 0);
            assemble_matrix_2d(*Q, *Q, A, B, c, K, 0, 0, map.data());
            co += static_cast<std::size_t>(nterms[g]) * np;
            mo += nmark[g];
            to += nterms[g];
        }
        std::copy(K.data.begin(), K.data.end(), data);
    });
}

int ref_assemble_vector_groups(void *sh, int ngroups, const int *kind, const int *degree, const int *nmark, const int *marks,
                               const int *all, const int *nterms, const int *a, const double *coeffs, double *f)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        for (int i = 0; i < V.dim(); ++i) f[i] = 0.0;
        std::size_t mo = 0, to = 0, co = 0;
        for (int g = 0; g < ngroups; ++g)
        {
            std::unique_ptr<Cache2D> Q = kind[g] == 0 ? std::make_unique<Cache2D>(V, degree[g])
                                                      : std::make_unique<Cache2D>(V, degree[g], std::vector<int>(marks + mo, marks + mo + nmark[g]), all[g] != 0);
            const std::size_t np = static_cast<std::size_t>(Q->nent()) * Q->nq();
            std::vector<std::vector<double>> c(static_cast<std::size_t>(nterms[g]));
            for (int t = 0; t < nterms[g]; ++t) c[t].assign(coeffs + co + t * np, coeffs + co + (t + 1) * np);
            std::vector<double> block(static_cast<std::size_t>(V.dim()), 0.0);
            assemble_vector_2d(*Q, std::vector<int>(a + to, a + to + nterms[g]), c, block);
            for (int i = 0; i < V.dim(); ++i) f[i] += block[i];
            co += static_cast<std::size_t>(nterms[g]) * np;
            mo += nmark[g];
            to += nterms[g];
        }
    });
}

int ref_integrand_eval(int ninstr, const int *code, int nreg, int nconst, const double *consts, long npts,
                       const double *xs, const double *ys, double *out)
{
    return guard([&] {
        Integrand P(std::vector<int>(code, code + 4 * ninstr), nreg, nconst, 0);
        P.evaluate(static_cast<std::size_t>(npts), xs, ys, nullptr, nullptr, {}, std::vector<double>(consts, consts + nconst), out);
    });
}

/// Integrand::evaluate with inputs: ninput arrays of npts values, one after another.
int ref_integrand_eval_in(int ninstr, const int *code, int nreg, int nconst, const double *consts, long npts, const double *xs,
                          const double *ys, int ninput, const double *inputs, double *out)
{
    return guard([&] {
        Integrand P(std::vector<int>(code, code + 4 * ninstr), nreg, nconst, ninput);
        std::vector<const double *> in(static_cast<std::size_t>(ninput));
        for (int i = 0; i < ninput; ++i) in[static_cast<std::size_t>(i)] = inputs + static_cast<std::size_t>(i) * npts;
        P.evaluate(static_cast<std::size_t>(npts), xs, ys, nullptr, nullptr, in, std::vector<double>(consts, consts + nconst), out);
    });
}

/// Cache2D::eval_multi: a known field (coefficients c, dim of them) at the cell points.
int ref_eval_field(void *sh, int degree, const double *c, int ncodes, const int *codes, double *out)
{
    return guard([&] {
        const Space2D &V = *static_cast<Space *>(sh)->V;
        Cache2D Q(V, degree);
        Q.eval_multi(c, false, codes, ncodes, out);
    });
}

int ref_integrand_eval_n(int ninstr, const int *code, int nreg, int nconst, const double *consts, long npts, const double *xs,
                         const double *ys, const double *nxs, const double *nys, double *out)
{
    return guard([&] {
        Integrand P(std::vector<int>(code, code + 4 * ninstr), nreg, nconst, 0);
        P.evaluate(static_cast<std::size_t>(npts), xs, ys, nxs, nys, {}, std::vector<double>(consts, consts + nconst), out);
    });
}

int ref_spmv(int n, const int *ip, const int *ix, const double *v, const double *x, double *y)
{
    return guard([&] { make_csr(n, n, ip, ix, v).apply(x, y); });
}

double ref_ddot(long n, const double *x, const double *y) { return ddot(x, y, static_cast<std::size_t>(n)); }

/// Preconditioner kinds: 0 none, 1 Jacobi, 2 ILU(0), 3 SSOR(omega).  ordering: 0 natural, 1 rcm.
/// The dispatch is python/src/femd_bindings.cpp:799-833: none and Jacobi take the fused cg_jacobi.
/// info: converged, iterations, breakdown; res: the true relative residual; min_pivot (ILU only).
int ref_cg(int n, const int *ip, const int *ix, const double *v, int pc, int ordering, double omega,
           const double *b, double *x, int max_iter, double tol, int *info, double *res)
{
    return guard([&] {
        CSRMatrix A = make_csr(n, n, ip, ix, v);
        auto Aop = [&A](const std::vector<double> &in, std::vector<double> &out) {
            out.resize(static_cast<std::size_t>(A.nrows()));
            A.apply(in.data(), out.data());
        };
        std::vector<double> xv(x, x + n), bv(b, b + n);
        const std::string ord = ordering ? "rcm" : "natural";
        krylov::CGResult r;
        if (pc == 0) r = krylov::cg_jacobi(Aop, nullptr, xv, bv, max_iter, tol);
        else if (pc == 1)
        {
            JacobiPreconditioner J(A);
            r = krylov::cg_jacobi(Aop, J.inverse_diagonal().data(), xv, bv, max_iter, tol);
        }
        else
        {
            std::unique_ptr<SparsePreconditioner> M;
            if (pc == 2) M = std::make_unique<ILU0Preconditioner>(A, ord);
            else M = std::make_unique<SSORPreconditioner>(A, omega, ord);
            auto Mop = [&M](const std::vector<double> &in, std::vector<double> &out) { (*M)(in, out); };
            r = krylov::cg(Aop, Mop, xv, bv, max_iter, tol);
        }
        std::copy(xv.begin(), xv.end(), x);
        info[0] = r.converged; info[1] = r.iterations; info[2] = r.breakdown;
        *res = r.residual;
    });
}

/// GMRES(restart) or LGMRES(restart, k_aug), left preconditioned, through krylov::solve.
/// res[0]: true residual, res[1]: preconditioned residual.  info: converged, iterations.
int ref_krylov(int n, const int *ip, const int *ix, const double *v, int pc, int ordering, double omega,
               const double *b, double *x, int lgmres, int restart, int max_iter, double tol, int k_aug, int *info, double *res)
{
    return guard([&] {
        CSRMatrix A = make_csr(n, n, ip, ix, v);
        auto Aop = [&A](const std::vector<double> &in, std::vector<double> &out) {
            out.resize(static_cast<std::size_t>(A.nrows()));
            A.apply(in.data(), out.data());
        };
        std::unique_ptr<SparsePreconditioner> M;
        const std::string ord = ordering ? "rcm" : "natural";
        if (pc == 1) M = std::make_unique<JacobiPreconditioner>(A);
        else if (pc == 2) M = std::make_unique<ILU0Preconditioner>(A, ord);
        else if (pc == 3) M = std::make_unique<SSORPreconditioner>(A, omega, ord);
        auto Mop = [&M](const std::vector<double> &in, std::vector<double> &out) {
            if (M) (*M)(in, out); else out = in;
        };
        std::vector<double> xv(x, x + n), bv(b, b + n);
        krylov::Result r = krylov::solve(Aop, Mop, xv, bv, lgmres ? krylov::Method::LGMRES : krylov::Method::GMRES,
                                         restart, max_iter, tol, k_aug);
        std::copy(xv.begin(), xv.end(), x);
        info[0] = r.converged; info[1] = r.iterations;
        res[0] = r.residual; res[1] = r.precond_residual;
    });
}

int ref_rcm(int n, const int *ip, const int *ix, int *perm)
{
    return guard([&] {
        std::vector<int> p = ordering::rcm_order(n, ip, ix);
        std::copy(p.begin(), p.end(), perm);
    });
}

/// z = M^{-1} r for the ILU(0) (kind 2), Jacobi (1) or SSOR (3) preconditioner of A.
/// min_pivot is written for ILU(0).
int ref_precond_apply(int n, const int *ip, const int *ix, const double *v, int kind, int ordering, double omega,
                      const double *r, double *z, double *min_pivot)
{
    return guard([&] {
        CSRMatrix A = make_csr(n, n, ip, ix, v);
        const std::string ord = ordering ? "rcm" : "natural";
        if (kind == 1) { JacobiPreconditioner J(A); J.apply(r, z); }
        else if (kind == 2)
        {
            ILU0Preconditioner P(A, ord);
            P.apply(r, z);
            if (min_pivot) *min_pivot = P.min_pivot();
        }
        else { SSORPreconditioner S(A, omega, ord); S.apply(r, z); }
    });
}

/// The asymmetry measure of csr.hpp:631-644.
double ref_asymmetry(int n, const int *ip, const int *ix, const double *v) { return asymmetry(make_csr(n, n, ip, ix, v)); }

} // extern "C"
