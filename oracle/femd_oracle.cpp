// femd_oracle.cpp -- the CPU oracle: a plain C++17 restatement of FEMd's 2D assemble -> SpMV ->
// Krylov path, in the reference's summation orders, with a C ABI (prefix orc_).
//
// TEST INFRASTRUCTURE ONLY.  The tests, __graft_entry__.smoke() and the cpu_baseline leg of
// bench.py load it as the checker of the HIP product (femd_amd/) and as the CPU baseline.
// The product never links, loads or calls it.
//
// Each function cites the reference file:line it restates (paths under the reference's
// include/femd/ and python/).  The arithmetic is the reference's, operation by operation
// (no fused multiply-add: built with -ffp-contract=off and no -march), so on the same inputs
// the results are bit-identical to the reference's serial build; the tests check that against
// the reference compiled in place (oracle/_ref) and against closed forms.
//
// Layout: a mesh and a space live in one Problem handle (orc_problem_*); the CSR routines
// take plain arrays.  Errors: -1 and orc_last_error().
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
template <class F> int guard(F &&f)
{
    try { f(); return 0; }
    catch (const std::exception &e) { g_err = e.what(); return -1; }
}

using I64 = std::int64_t;
inline I64 ekey(int a, int b)
{
    if (a > b) std::swap(a, b);
    return (static_cast<I64>(a) << 32) | static_cast<std::uint32_t>(b);
}

// ---------------------------------------------------------------------------------- mesh
// mesh/structured.hpp:99-221 (mesh_from_arrays), 260-351 (structured_grid_mesh, rectangle_mesh)
struct Mesh {
    int nv = 3;                             // vertices per cell: 3 triangles, 4 quadrilaterals
    std::vector<double> xy;                 // np x 2
    std::vector<int> tri;                   // nt x nv, counter-clockwise
    std::vector<int> nbr;                   // nt x nv, across local edge i
    std::vector<int> seg;                   // ns x 2
    std::vector<int> smark;                 // ns
    std::unordered_map<I64, int> segidx;    // edge key -> segment index
    int np() const { return static_cast<int>(xy.size() / 2); }
    int nt() const { return static_cast<int>(tri.size() / nv); }
    // the vertices of local edge i of cell c: opposite vertex i on a triangle (mesh2d.hpp),
    // from vertex i to vertex i+1 on a quadrilateral (quad_mesh.hpp:4-9)
    void edge(int c, int i, int &a, int &b) const
    {
        const int *t = &tri[static_cast<std::size_t>(c) * nv];
        if (nv == 3) { a = t[(i + 1) % 3]; b = t[(i + 2) % 3]; }
        else { a = t[i]; b = t[(i + 1) % 4]; }
    }
    int segment_marker(int a, int b) const
    {
        auto it = segidx.find(ekey(a, b));
        return it == segidx.end() ? -1 : smark[it->second];
    }
    void add_segment(int a, int b, int m)
    {
        auto it = segidx.find(ekey(a, b));
        if (it != segidx.end()) { smark[it->second] = m; return; }
        segidx[ekey(a, b)] = static_cast<int>(smark.size());
        seg.push_back(a); seg.push_back(b); smark.push_back(m);
    }
};

// structured.hpp:99-221.  The orientation test is the sign of the plain determinant: the
// reference uses exact predicates, which agree with it on every non-degenerate triangle.
Mesh build_mesh(std::vector<double> xy, std::vector<int> tri, const std::vector<int> &segs,
                const std::vector<int> &smarks, int marker)
{
    Mesh M;
    const int np = static_cast<int>(xy.size() / 2), nt = static_cast<int>(tri.size() / 3);
    if (np < 3 || nt < 1) throw std::invalid_argument("mesh: need at least 3 points and 1 triangle");
    for (int t = 0; t < nt; ++t)
    {
        int *T = &tri[3 * t];
        for (int k = 0; k < 3; ++k)
            if (T[k] < 0 || T[k] >= np) throw std::invalid_argument("mesh: vertex index out of range");
        const double ax = xy[2 * T[0]], ay = xy[2 * T[0] + 1], bx = xy[2 * T[1]], by = xy[2 * T[1] + 1];
        const double cx = xy[2 * T[2]], cy = xy[2 * T[2] + 1];
        const double det = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
        if (det == 0.0) throw std::invalid_argument("mesh: degenerate triangle");
        if (det < 0.0) std::swap(T[1], T[2]);
    }
    M.xy = std::move(xy);
    M.tri = std::move(tri);
    M.nbr.assign(static_cast<std::size_t>(nt) * 3, -1);
    std::unordered_map<I64, std::pair<int, int>> first;     // edge -> (t, i) of its first side
    first.reserve(static_cast<std::size_t>(3 * nt));
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < 3; ++i)
        {
            const int a = M.tri[3 * t + (i + 1) % 3], b = M.tri[3 * t + (i + 2) % 3];
            auto it = first.find(ekey(a, b));
            if (it == first.end()) { first.emplace(ekey(a, b), std::make_pair(t, i)); continue; }
            const int u = it->second.first, j = it->second.second;
            if (M.nbr[3 * u + j] >= 0) throw std::invalid_argument("mesh: edge shared by more than two triangles");
            M.nbr[3 * t + i] = u; M.nbr[3 * u + j] = t;
        }
    // boundary edges in triangle order, directed as in their triangle
    std::vector<std::vector<int>> out(static_cast<std::size_t>(np));
    std::vector<std::array<int, 2>> bedge;
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < 3; ++i)
            if (M.nbr[3 * t + i] < 0)
            {
                const int a = M.tri[3 * t + (i + 1) % 3], b = M.tri[3 * t + (i + 2) % 3];
                out[a].push_back(static_cast<int>(bedge.size()));
                bedge.push_back({a, b});
            }
    const int ns = static_cast<int>(segs.size() / 2);
    for (int s = 0; s < ns; ++s)
    {
        if (first.find(ekey(segs[2 * s], segs[2 * s + 1])) == first.end())
            throw std::invalid_argument("mesh: a segment is not an edge");
        M.add_segment(segs[2 * s], segs[2 * s + 1], smarks.empty() ? marker : smarks[s]);
    }
    // unlisted boundary edges, walked loop by loop from the lowest vertex (structured.hpp:182-200)
    std::vector<char> done(bedge.size(), 0);
    for (int start = 0; start < np; ++start)
        for (int e0 : out[start])
        {
            int e = e0;
            while (e >= 0 && !done[e])
            {
                done[e] = 1;
                const int a = bedge[e][0], b = bedge[e][1];
                if (M.segidx.find(ekey(a, b)) == M.segidx.end()) M.add_segment(a, b, marker);
                int next = -1;
                for (int f : out[b]) if (!done[f]) { next = f; break; }
                e = next;
            }
        }
    return M;
}

// structured.hpp:260-351 (the grid, the diagonal patterns, segments by side)
Mesh rectangle(double x0, double y0, double x1, double y1, int nx, int ny, int diag, const int markers[4])
{
    if (!(x1 > x0 && y1 > y0) || nx < 1 || ny < 1) throw std::invalid_argument("rectangle: bad extents or counts");
    std::vector<double> g(static_cast<std::size_t>(nx + 1) * (ny + 1) * 2);
    for (int j = 0; j <= ny; ++j)
    {
        const double y = j == ny ? y1 : y0 + (y1 - y0) * j / ny;
        for (int i = 0; i <= nx; ++i)
        {
            g[2 * (j * (nx + 1) + i)] = i == nx ? x1 : x0 + (x1 - x0) * i / nx;
            g[2 * (j * (nx + 1) + i) + 1] = y;
        }
    }
    auto V = [nx](int i, int j) { return j * (nx + 1) + i; };
    const int nc0 = (nx + 1) * (ny + 1);
    if (diag == 3)                                        // crossed: one centre per cell
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
            {
                const int a = V(i, j), b = V(i + 1, j), c = V(i + 1, j + 1), d = V(i, j + 1);
                g.push_back(0.25 * (g[2 * a] + g[2 * b] + g[2 * c] + g[2 * d]));
                g.push_back(0.25 * (g[2 * a + 1] + g[2 * b + 1] + g[2 * c + 1] + g[2 * d + 1]));
            }
    std::vector<int> tri;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
            const int a = V(i, j), b = V(i + 1, j), c = V(i + 1, j + 1), d = V(i, j + 1);
            const bool right = diag == 0 || (diag == 2 && (i + j) % 2 == 0);
            if (diag == 3)
            {
                const int m = nc0 + j * nx + i;
                const int t4[12] = {a, b, m, b, c, m, c, d, m, d, a, m};
                tri.insert(tri.end(), t4, t4 + 12);
            }
            else if (right) { const int t2[6] = {a, b, c, a, c, d}; tri.insert(tri.end(), t2, t2 + 6); }
            else { const int t2[6] = {a, b, d, b, c, d}; tri.insert(tri.end(), t2, t2 + 6); }
        }
    std::vector<int> segs, smk;
    for (int i = 0; i < nx; ++i) { segs.push_back(V(i, 0)); segs.push_back(V(i + 1, 0)); smk.push_back(markers[0]); }
    for (int j = 0; j < ny; ++j) { segs.push_back(V(nx, j)); segs.push_back(V(nx, j + 1)); smk.push_back(markers[1]); }
    for (int i = nx; i > 0; --i) { segs.push_back(V(i, ny)); segs.push_back(V(i - 1, ny)); smk.push_back(markers[2]); }
    for (int j = ny; j > 0; --j) { segs.push_back(V(0, j)); segs.push_back(V(0, j - 1)); smk.push_back(markers[3]); }
    return build_mesh(std::move(g), std::move(tri), segs, smk, markers[0]);
}

// mesh/quad_mesh.hpp:50-103 (the QuadMesh constructor from arrays): clockwise quads re-oriented
// by swapping vertices 1 and 3, strict convexity required, neighbours across each edge, the
// listed segments first, then every other boundary edge (in quad order) with `marker`.  The
// convexity test is the sign of the plain cross product (the reference's exact predicate agrees
// on every quad that is not nearly degenerate).
Mesh build_quad_mesh(std::vector<double> xy, std::vector<int> q, const std::vector<int> &segs, const std::vector<int> &smarks,
                     int marker)
{
    Mesh M;
    M.nv = 4;
    const int np = static_cast<int>(xy.size() / 2), nq = static_cast<int>(q.size() / 4);
    auto P = [&](int v, int d) { return xy[2 * static_cast<std::size_t>(v) + d]; };
    for (int t = 0; t < nq; ++t)
    {
        int *Q = &q[4 * static_cast<std::size_t>(t)];
        for (int k = 0; k < 4; ++k)
            if (Q[k] < 0 || Q[k] >= np) throw std::invalid_argument("quad mesh: vertex index out of range");
        double s = 0.0;
        for (int i = 0; i < 4; ++i) s += P(Q[i], 0) * P(Q[(i + 1) % 4], 1) - P(Q[(i + 1) % 4], 0) * P(Q[i], 1);
        if (0.5 * s < 0.0) std::swap(Q[1], Q[3]);
        for (int i = 0; i < 4; ++i)
        {
            const int a = Q[(i + 3) % 4], b = Q[i], c = Q[(i + 1) % 4];
            const double o = (P(b, 0) - P(a, 0)) * (P(c, 1) - P(a, 1)) - (P(b, 1) - P(a, 1)) * (P(c, 0) - P(a, 0));
            if (!(o > 0.0)) throw std::invalid_argument("quad mesh: a quad is degenerate or not convex");
        }
    }
    M.xy = std::move(xy);
    M.tri = std::move(q);
    M.nbr.assign(static_cast<std::size_t>(nq) * 4, -1);
    std::unordered_map<I64, std::pair<int, int>> first;
    for (int t = 0; t < nq; ++t)
        for (int i = 0; i < 4; ++i)
        {
            int a, b;
            M.edge(t, i, a, b);
            auto it = first.find(ekey(a, b));
            if (it == first.end()) { first.emplace(ekey(a, b), std::make_pair(t, i)); continue; }
            const int u = it->second.first, j = it->second.second;
            if (M.nbr[4 * u + j] >= 0) throw std::invalid_argument("quad mesh: edge shared by more than two quads");
            M.nbr[4 * t + i] = u; M.nbr[4 * u + j] = t;
        }
    const int ns = static_cast<int>(segs.size() / 2);
    for (int s2 = 0; s2 < ns; ++s2)
    {
        if (first.find(ekey(segs[2 * s2], segs[2 * s2 + 1])) == first.end())
            throw std::invalid_argument("quad mesh: a segment is not an edge");
        M.add_segment(segs[2 * s2], segs[2 * s2 + 1], smarks.empty() ? marker : smarks[s2]);
    }
    for (int t = 0; t < nq; ++t)
        for (int i = 0; i < 4; ++i)
            if (M.nbr[4 * t + i] < 0)
            {
                int a, b;
                M.edge(t, i, a, b);
                if (M.segidx.find(ekey(a, b)) == M.segidx.end()) M.add_segment(a, b, marker);
            }
    return M;
}

// python/femd/quadmesh.py:146-170 (rectangle_quad_mesh): the grid and the sides of
// rectangle_mesh, quad (i, j) = j nx + i with vertices (i, j), (i+1, j), (i+1, j+1), (i, j+1)
Mesh quad_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, const int markers[4])
{
    const Mesh T = rectangle(x0, y0, x1, y1, nx, ny, 0, markers);
    std::vector<double> g(T.xy.begin(), T.xy.begin() + 2 * static_cast<std::size_t>(nx + 1) * (ny + 1));
    std::vector<int> q;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
            const int v = j * (nx + 1) + i;
            const int c4[4] = {v, v + 1, v + nx + 2, v + nx + 1};
            q.insert(q.end(), c4, c4 + 4);
        }
    return build_quad_mesh(std::move(g), std::move(q), T.seg, T.smark, markers[0]);
}

// ------------------------------------------------------------------------ reference element
// fe/simplex_basis.hpp:38-55 (orthonormal Jacobi), 77-114 (Dubiner modes), 117-148 (inverse),
// 202-296 (ReferenceTriangle; equispaced or warp-and-blend nodes).
void jacobi_p(double x, double al, double be, int N, double *P)
{
    const double ab = al + be;
    const double g0 = std::pow(2.0, ab + 1.0) / (ab + 1.0) * std::tgamma(al + 1.0) * std::tgamma(be + 1.0) / std::tgamma(ab + 1.0);
    P[0] = 1.0 / std::sqrt(g0);
    if (N == 0) return;
    const double g1 = (al + 1.0) * (be + 1.0) / (ab + 3.0) * g0;
    P[1] = ((ab + 2.0) * x / 2.0 + (al - be) / 2.0) / std::sqrt(g1);
    double aold = 2.0 / (2.0 + ab) * std::sqrt((al + 1.0) * (be + 1.0) / (ab + 3.0));
    for (int i = 1; i < N; ++i)
    {
        const double h1 = 2.0 * i + ab;
        const double anew = 2.0 / (h1 + 2.0) * std::sqrt((i + 1.0) * (i + 1.0 + ab) * (i + 1.0 + al) * (i + 1.0 + be) / (h1 + 1.0) / (h1 + 3.0));
        const double bnew = -(al * al - be * be) / h1 / (h1 + 2.0);
        P[i + 1] = 1.0 / anew * (-aold * P[i - 1] + (x - bnew) * P[i]);
        aold = anew;
    }
}

double jacobi_d(double x, double al, double be, int n)
{
    if (n == 0) return 0.0;
    std::vector<double> P(static_cast<std::size_t>(n));
    jacobi_p(x, al + 1.0, be + 1.0, n - 1, P.data());
    return std::sqrt(n * (n + al + be + 1.0)) * P[n - 1];
}

void dubiner_modes(int k, double xi, double eta, double *v, double *dxi, double *deta)
{
    const double r = 2.0 * xi - 1.0, s = 2.0 * eta - 1.0;
    const double a = (std::abs(1.0 - s) > 1e-14) ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0;
    const double b = s;
    std::vector<double> Pa(static_cast<std::size_t>(k) + 1), dPa(static_cast<std::size_t>(k) + 1), Pb(static_cast<std::size_t>(k) + 1);
    jacobi_p(a, 0.0, 0.0, k, Pa.data());
    for (int i = 0; i <= k; ++i) dPa[i] = jacobi_d(a, 0.0, 0.0, i);
    const double hb = 0.5 * (1.0 - b);
    int m = 0;
    for (int i = 0; i <= k; ++i)
    {
        jacobi_p(b, 2.0 * i + 1.0, 0.0, k - i, Pb.data());
        const double hbi = std::pow(hb, i), hbim1 = i > 0 ? std::pow(hb, i - 1) : 0.0;
        const double c = std::pow(2.0, i + 0.5);
        for (int j = 0; j <= k - i; ++j, ++m)
        {
            const double fa = Pa[i], gb = Pb[j];
            v[m] = c * fa * gb * hbi;
            if (!dxi) continue;
            const double dfa = dPa[i], dgb = jacobi_d(b, 2.0 * i + 1.0, 0.0, j);
            double dr = dfa * gb;
            if (i > 0) dr *= hbim1;
            double ds = dfa * gb * 0.5 * (1.0 + a);
            if (i > 0) ds *= hbim1;
            double tmp = dgb * hbi;
            if (i > 0) tmp -= 0.5 * i * gb * hbim1;
            ds += fa * tmp;
            dr *= c; ds *= c;
            dxi[m] = 2.0 * dr;
            deta[m] = 2.0 * ds;
        }
    }
}

std::vector<double> gauss_jordan_inverse(std::vector<double> A, int n)
{
    std::vector<double> X(static_cast<std::size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) X[static_cast<std::size_t>(i) * n + i] = 1.0;
    for (int c = 0; c < n; ++c)
    {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(A[r * n + c]) > std::abs(A[piv * n + c])) piv = r;
        if (A[piv * n + c] == 0.0) throw std::runtime_error("singular Vandermonde matrix");
        if (piv != c)
            for (int j = 0; j < n; ++j) { std::swap(A[piv * n + j], A[c * n + j]); std::swap(X[piv * n + j], X[c * n + j]); }
        const double d = A[c * n + c];
        for (int j = 0; j < n; ++j) { A[c * n + j] /= d; X[c * n + j] /= d; }
        for (int r = 0; r < n; ++r)
        {
            if (r == c) continue;
            const double f = A[r * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) { A[r * n + j] -= f * A[c * n + j]; X[r * n + j] -= f * X[c * n + j]; }
        }
    }
    return X;
}

// fe/dg_space.hpp:38-58: P_n^(m)(x), n = 0..p, m = 0..nder, by
// n P_n^(m) = (2n-1)(x P_{n-1}^(m) + m P_{n-1}^(m-1)) - (n-1) P_{n-2}^(m); out[m (p+1) + n].
void legendre_rows(int p, int nder, double x, double *out)
{
    const int nl = p + 1;
    for (int m = 0; m <= nder; ++m)
        for (int n = 0; n <= p; ++n)
        {
            if (n == 0) { out[m * nl] = m == 0 ? 1.0 : 0.0; continue; }
            const double a = out[m * nl + n - 1];
            const double b = m > 0 ? out[(m - 1) * nl + n - 1] : 0.0;
            const double c = n >= 2 ? out[m * nl + n - 2] : 0.0;
            out[m * nl + n] = ((2 * n - 1) * (x * a + m * b) - (n - 1) * c) / n;
        }
}

// fe/dg_space.hpp:60-80: the p+1 Gauss-Lobatto points on [-1, 1], the interior ones by Newton on
// P_p' / P_p'' from -cos(pi i / p), at most 100 steps, stopping once a step is below 1e-16.
std::vector<double> lobatto_points(int p)
{
    std::vector<double> x(static_cast<std::size_t>(p) + 1), t(3 * (static_cast<std::size_t>(p) + 1));
    x[0] = -1.0; x[p] = 1.0;
    for (int i = 1; i < p; ++i)
    {
        double z = -std::cos(3.14159265358979323846 * i / p);
        for (int it = 0; it < 100; ++it)
        {
            legendre_rows(p, 2, z, t.data());
            const double d = t[(p + 1) + p] / t[2 * (p + 1) + p];
            z -= d;
            if (std::fabs(d) < 1e-16) break;
        }
        x[i] = z;
    }
    return x;
}

// simplex_basis.hpp:150-166: Hesthaven's warp, the Lobatto-minus-uniform displacement
// interpolated at r (uniform nodes) and divided by 1 - r^2 (zero at the ends).
double warp_of(int N, double r, const std::vector<double> &lgl)
{
    double w = 0.0;
    for (int i = 0; i <= N; ++i)
    {
        const double ri = -1.0 + 2.0 * i / N;
        double L = 1.0;
        for (int j = 0; j <= N; ++j)
            if (j != i) L *= (r - (-1.0 + 2.0 * j / N)) / (ri - (-1.0 + 2.0 * j / N));
        w += L * (lgl[i] - ri);
    }
    return std::abs(r) < 1.0 - 1e-10 ? w / (1.0 - r * r) : 0.0;
}

// simplex_basis.hpp:168-193: Warburton's warp-and-blend points of degree N on the equilateral
// triangle (blend optimum alpha per degree, 5/3 above 15), mapped back to (xi, eta), lattice order.
void warp_blend_points(int N, std::vector<double> &px, std::vector<double> &py)
{
    static const double aopt[15] = {0.0, 0.0, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                    1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
    const double al = N < 16 ? aopt[N - 1] : 5.0 / 3.0;
    const std::vector<double> lgl = lobatto_points(N);
    const double pi = 3.14159265358979323846, s3 = std::sqrt(3.0);
    for (int n = 0; n <= N; ++n)
        for (int m = 0; m <= N - n; ++m)
        {
            const double L1 = static_cast<double>(n) / N, L3 = static_cast<double>(m) / N, L2 = 1.0 - L1 - L3;
            double x = -L2 + L3, y = (-L2 - L3 + 2.0 * L1) / s3;
            const double w1 = 4.0 * L2 * L3 * warp_of(N, L3 - L2, lgl) * (1.0 + (al * L1) * (al * L1));
            const double w2 = 4.0 * L1 * L3 * warp_of(N, L1 - L3, lgl) * (1.0 + (al * L2) * (al * L2));
            const double w3 = 4.0 * L1 * L2 * warp_of(N, L2 - L1, lgl) * (1.0 + (al * L3) * (al * L3));
            x += w1 + std::cos(2.0 * pi / 3.0) * w2 + std::cos(4.0 * pi / 3.0) * w3;
            y += std::sin(2.0 * pi / 3.0) * w2 + std::sin(4.0 * pi / 3.0) * w3;
            const double l1 = (s3 * y + 1.0) / 3.0, l2 = (-3.0 * x - s3 * y + 2.0) / 6.0, l3 = (3.0 * x - s3 * y + 2.0) / 6.0;
            const double r = -l2 + l3 - l1, s = -l2 - l3 + l1;
            px.push_back(0.5 * (r + 1.0));
            py.push_back(0.5 * (s + 1.0));
        }
}

// simplex_basis.hpp:253-285: equispaced nodes, or (lobatto) symmetrized Gauss-Lobatto edge nodes
// and the warp-and-blend points strictly inside the triangle.
struct RefElement {
    int k = 1, n = 3;
    bool equi = true;
    std::vector<double> nx, ny, Vinv;
    explicit RefElement(int deg, bool equispaced = true) : k(deg), n((deg + 1) * (deg + 2) / 2), equi(equispaced)
    {
        const double vx[3] = {0.0, 1.0, 0.0}, vy[3] = {0.0, 0.0, 1.0};
        for (int v = 0; v < 3; ++v) { nx.push_back(vx[v]); ny.push_back(vy[v]); }
        std::vector<double> z;
        if (equi) for (int j = 0; j <= k; ++j) z.push_back(-1.0 + 2.0 * j / k);
        else z = lobatto_points(k);
        for (int j = 0; j <= k; ++j) z[j] = 0.5 * (z[j] - z[k - j]);     // simplex_basis.hpp:259, symmetrized
        for (int e = 0; e < 3; ++e)
        {
            const int a = (e + 1) % 3, b = (e + 2) % 3;
            for (int j = 1; j < k; ++j)
            {
                const double t = 0.5 * (1.0 + z[j]);
                nx.push_back((1.0 - t) * vx[a] + t * vx[b]);
                ny.push_back((1.0 - t) * vy[a] + t * vy[b]);
            }
        }
        if (equi)
        {
            for (int j = 1; j < k; ++j)
                for (int i = 1; i + j < k; ++i) { nx.push_back(static_cast<double>(i) / k); ny.push_back(static_cast<double>(j) / k); }
        }
        else if (k >= 3)
        {
            std::vector<double> px, py;
            warp_blend_points(k, px, py);
            for (std::size_t p = 0; p < px.size(); ++p)
                if (px[p] > 1e-10 && py[p] > 1e-10 && 1.0 - px[p] - py[p] > 1e-10) { nx.push_back(px[p]); ny.push_back(py[p]); }
        }
        if (static_cast<int>(nx.size()) != n) throw std::logic_error("reference element: node count mismatch");
        std::vector<double> V(static_cast<std::size_t>(n) * n), psi(static_cast<std::size_t>(n));
        for (int i = 0; i < n; ++i)
        {
            dubiner_modes(k, nx[i], ny[i], psi.data(), nullptr, nullptr);
            for (int m = 0; m < n; ++m) V[i * n + m] = psi[m];
        }
        Vinv = gauss_jordan_inverse(V, n);
    }
    // simplex_basis.hpp:235-252: out[r*n + l] = sum_m psi_r[m] Vinv[m][l], zero modes skipped
    void eval(double xi, double eta, double *out) const
    {
        std::vector<double> psi(static_cast<std::size_t>(3) * n);
        dubiner_modes(k, xi, eta, psi.data(), psi.data() + n, psi.data() + 2 * n);
        for (int r = 0; r < 3; ++r)
        {
            const double *pr = psi.data() + r * n;
            double *o = out + r * n;
            for (int l = 0; l < n; ++l) o[l] = 0.0;
            for (int m = 0; m < n; ++m)
            {
                const double a = pr[m];
                if (a == 0.0) continue;
                for (int l = 0; l < n; ++l) o[l] += a * Vinv[m * n + l];
            }
        }
    }
};

// fe/lagrange_space_quad.hpp:28-110 (ReferenceQuad, equispaced or Gauss-Lobatto nodes): the tensor product of
// 1D Lagrange polynomials in barycentric form on k+1 symmetrized points; local order vertices,
// edge nodes along each edge, interior row by row.
struct RefQuad {
    int k = 1, n = 4;
    std::vector<double> t, w, nx, ny;
    std::vector<int> ii, jj;
    explicit RefQuad(int deg, bool equispaced = true) : k(deg), n((deg + 1) * (deg + 1))
    {
        std::vector<double> z;
        if (equispaced) for (int j = 0; j <= k; ++j) z.push_back(-1.0 + 2.0 * j / k);
        else z = lobatto_points(k);
        t.resize(static_cast<std::size_t>(k) + 1);
        for (int j = 0; j <= k; ++j) t[j] = 0.5 * (1.0 + 0.5 * (z[j] - z[k - j]));
        t[0] = 0.0; t[k] = 1.0;
        w.assign(static_cast<std::size_t>(k) + 1, 1.0);
        for (int j = 0; j <= k; ++j)
            for (int m = 0; m <= k; ++m) if (m != j) w[j] /= (t[j] - t[m]);
        auto add = [&](int i, int j) { ii.push_back(i); jj.push_back(j); };
        add(0, 0); add(k, 0); add(k, k); add(0, k);
        for (int j = 1; j < k; ++j) add(j, 0);
        for (int j = 1; j < k; ++j) add(k, j);
        for (int j = 1; j < k; ++j) add(k - j, k);
        for (int j = 1; j < k; ++j) add(0, k - j);
        for (int j = 1; j < k; ++j)
            for (int i = 1; i < k; ++i) add(i, j);
        for (int l = 0; l < n; ++l) { nx.push_back(t[ii[l]]); ny.push_back(t[jj[l]]); }
    }
    void basis_1d(double x, double *L, double *dL) const
    {
        for (int j = 0; j <= k; ++j)
        {
            double p = 1.0, dp = 0.0;
            for (int m = 0; m <= k; ++m)
            {
                if (m == j) continue;
                dp = dp * (x - t[m]) + p;
                p *= (x - t[m]);
            }
            L[j] = w[j] * p;
            dL[j] = w[j] * dp;
        }
    }
    void eval(double xi, double eta, double *out) const
    {
        std::vector<double> a(static_cast<std::size_t>(4) * (k + 1));
        double *Lx = a.data(), *dLx = Lx + (k + 1), *Ly = dLx + (k + 1), *dLy = Ly + (k + 1);
        basis_1d(xi, Lx, dLx);
        basis_1d(eta, Ly, dLy);
        for (int l = 0; l < n; ++l)
        {
            out[l] = Lx[ii[l]] * Ly[jj[l]];
            out[n + l] = dLx[ii[l]] * Ly[jj[l]];
            out[2 * n + l] = Lx[ii[l]] * dLy[jj[l]];
        }
    }
};

// ------------------------------------------------------------------------------ quadrature
// quadrature/gauss_legendre.hpp:30-137 and quadrature/triangle.hpp:22-57
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(static_cast<std::size_t>(n), 0.0); w.assign(static_cast<std::size_t>(n), 0.0);
    if (n == 1) { x = {0.0}; w = {2.0}; }
    else if (n == 2) { x = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)}; w = {1.0, 1.0}; }
    else if (n == 3) { x = {-std::sqrt(3.0 / 5.0), 0.0, std::sqrt(3.0 / 5.0)}; w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0}; }
    else if (n == 4)
    {
        const double a = std::sqrt((3.0 + 2.0 * std::sqrt(6.0 / 5.0)) / 7.0), b = std::sqrt((3.0 - 2.0 * std::sqrt(6.0 / 5.0)) / 7.0);
        const double wa = (18.0 - std::sqrt(30.0)) / 36.0, wb = (18.0 + std::sqrt(30.0)) / 36.0;
        x = {-a, -b, b, a}; w = {wa, wb, wb, wa};
    }
    else if (n == 5)
    {
        const double a = std::sqrt(5.0 + 2.0 * std::sqrt(10.0 / 7.0)) / 3.0, b = std::sqrt(5.0 - 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        const double wa = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0, wb = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        x = {-a, -b, 0.0, b, a}; w = {wa, wb, 128.0 / 225.0, wb, wa};
    }
    else
    {
        const double PI = 3.14159265358979323846;
        for (int i = 0; i < n; ++i)
        {
            double z = std::cos(PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it)
            {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= n; ++k) { const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k; p0 = p1; p1 = p2; }
                const double dp = n * (z * p1 - p0) / (z * z - 1.0), dz = -p1 / dp;
                z += dz;
                if (std::abs(dz) <= 1e-15) break;
            }
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) { const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k; p0 = p1; p1 = p2; }
            const double dp = n * (z * p1 - p0) / (z * z - 1.0);
            x[i] = z; w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
        }
    }
}

void triangle_rule(int d, std::vector<double> &qx, std::vector<double> &qy, std::vector<double> &qw)
{
    qx.clear(); qy.clear(); qw.clear();
    auto add = [&](double a, double b, double c) { qx.push_back(a); qy.push_back(b); qw.push_back(c); };
    if (d < 0) d = 0;
    if (d <= 1) add(1.0 / 3.0, 1.0 / 3.0, 0.5);
    else if (d == 2) { add(1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0); add(2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0); add(1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0); }
    else if (d <= 5)
    {
        const double s15 = std::sqrt(15.0);
        add(1.0 / 3.0, 1.0 / 3.0, 9.0 / 80.0);
        const double a1 = (6.0 - s15) / 21.0, w1 = (155.0 - s15) / 2400.0, a2 = (6.0 + s15) / 21.0, w2 = (155.0 + s15) / 2400.0;
        add(a1, a1, w1); add(1.0 - 2.0 * a1, a1, w1); add(a1, 1.0 - 2.0 * a1, w1);
        add(a2, a2, w2); add(1.0 - 2.0 * a2, a2, w2); add(a2, 1.0 - 2.0 * a2, w2);
    }
    else
    {
        const int nu = (d + 2) / 2, nv = (d + 3) / 2;
        std::vector<double> ux, uw, vx, vw;
        gauss_legendre(nu, ux, uw); gauss_legendre(nv, vx, vw);
        for (int j = 0; j < nv; ++j)
        {
            const double v = 0.5 * (vx[j] + 1.0), wv = 0.5 * vw[j];
            for (int i = 0; i < nu; ++i)
            {
                const double u = 0.5 * (ux[i] + 1.0), wu = 0.5 * uw[i];
                add(u * (1.0 - v), v, wu * wv * (1.0 - v));
            }
        }
    }
}

// fe/cache_2d.hpp:75-84: the tensor Gauss-Legendre rule with max(1, (d+3)/2) points per direction
// on [0,1]^2, xi running fastest
void quad_rule(int d, std::vector<double> &qx, std::vector<double> &qy, std::vector<double> &qw)
{
    qx.clear(); qy.clear(); qw.clear();
    std::vector<double> g, gw;
    gauss_legendre(std::max(1, (d + 3) / 2), g, gw);
    const int m = static_cast<int>(g.size());
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i)
        {
            qx.push_back(0.5 * (g[i] + 1.0));
            qy.push_back(0.5 * (g[j] + 1.0));
            qw.push_back(0.25 * gw[i] * gw[j]);
        }
}

// --------------------------------------------------------------------------------- space
// fe/space_2d.hpp:390-526 (edges in first-met order, raw numbering, Dirichlet elimination,
// adapted numbering) and fe/lagrange_space_2d.hpp:32-67 (per-cell J, J^{-1}, det).
struct Problem {
    Mesh mesh;
    int k = 1, nloc = 3, nedges = 0, raw_dim = 0, dim = 0;
    std::vector<int> cell_raw, cell_dof, ad2raw, raw2ad;
    std::vector<double> raw_xy, J, Ji, det;
    std::unique_ptr<RefElement> ref;
    std::unique_ptr<RefQuad> refq;            // quadrilateral meshes
    bool quad() const { return mesh.nv == 4; }
    void ref_eval(double xi, double eta, double *out) const { if (refq) refq->eval(xi, eta, out); else ref->eval(xi, eta, out); }
    double node_xi(int l) const { return refq ? refq->nx[l] : ref->nx[l]; }
    double node_eta(int l) const { return refq ? refq->ny[l] : ref->ny[l]; }
};

// fe/lagrange_space_quad.hpp:155-172: the bilinear map of quad c and its Jacobian
// [dx/dxi, dx/deta, dy/dxi, dy/deta]
void quad_map(const Mesh &M, int c, double xi, double eta, double &x, double &y, double *J)
{
    const int *q = &M.tri[4 * static_cast<std::size_t>(c)];
    const double *P0 = &M.xy[2 * q[0]], *P1 = &M.xy[2 * q[1]], *P2 = &M.xy[2 * q[2]], *P3 = &M.xy[2 * q[3]];
    const double a0 = (1 - xi) * (1 - eta), a1 = xi * (1 - eta), a2 = xi * eta, a3 = (1 - xi) * eta;
    x = a0 * P0[0] + a1 * P1[0] + a2 * P2[0] + a3 * P3[0];
    y = a0 * P0[1] + a1 * P1[1] + a2 * P2[1] + a3 * P3[1];
    if (J)
    {
        J[0] = (P1[0] - P0[0]) * (1 - eta) + (P2[0] - P3[0]) * eta;
        J[1] = (P3[0] - P0[0]) * (1 - xi) + (P2[0] - P1[0]) * xi;
        J[2] = (P1[1] - P0[1]) * (1 - eta) + (P2[1] - P3[1]) * eta;
        J[3] = (P3[1] - P0[1]) * (1 - xi) + (P2[1] - P1[1]) * xi;
    }
}

void build_space(Problem &P, int k, const std::vector<int> &dirichlet_in, bool all_exterior, bool equispaced)
{
    if (k < 1) throw std::invalid_argument("space: degree must be >= 1");
    std::vector<int> dirichlet = dirichlet_in;
    std::sort(dirichlet.begin(), dirichlet.end());
    dirichlet.erase(std::unique(dirichlet.begin(), dirichlet.end()), dirichlet.end());
    const Mesh &M = P.mesh;
    const int nc = M.nt(), np = M.np(), nv = M.nv;
    P.k = k;
    if (nv == 4) P.refq = std::make_unique<RefQuad>(k, equispaced);
    else P.ref = std::make_unique<RefElement>(k, equispaced);
    P.nloc = nv == 4 ? P.refq->n : P.ref->n;
    const int n = P.nloc, de = k - 1, nint = nv == 4 ? (k - 1) * (k - 1) : (k - 1) * (k - 2) / 2, f0 = nv + nv * (k - 1);
    P.J.resize(4 * static_cast<std::size_t>(nc)); P.Ji.resize(P.J.size()); P.det.resize(nc);
    for (int c = 0; c < nc && nv == 3; ++c)
    {
        const int *t = &M.tri[3 * c];
        double *J = &P.J[4 * c], *Ji = &P.Ji[4 * c];
        J[0] = M.xy[2 * t[1]] - M.xy[2 * t[0]]; J[1] = M.xy[2 * t[2]] - M.xy[2 * t[0]];
        J[2] = M.xy[2 * t[1] + 1] - M.xy[2 * t[0] + 1]; J[3] = M.xy[2 * t[2] + 1] - M.xy[2 * t[0] + 1];
        const double d = J[0] * J[3] - J[1] * J[2];
        if (!(d > 0.0)) throw std::invalid_argument("space: degenerate or clockwise cell");
        P.det[c] = d;
        Ji[0] = J[3] / d; Ji[1] = -J[1] / d; Ji[2] = -J[2] / d; Ji[3] = J[0] / d;
    }
    // edges, first met (cell c, local edge i: Mesh::edge)
    std::vector<int> cedge(nv * static_cast<std::size_t>(nc)), cflip(nv * static_cast<std::size_t>(nc));
    std::vector<std::array<int, 2>> edges;
    std::unordered_map<I64, int> emap;
    emap.reserve(2 * static_cast<std::size_t>(nc));
    for (int c = 0; c < nc; ++c)
        for (int i = 0; i < nv; ++i)
        {
            int a, b;
            M.edge(c, i, a, b);
            auto it = emap.find(ekey(a, b));
            int g;
            if (it == emap.end()) { g = static_cast<int>(edges.size()); edges.push_back({std::min(a, b), std::max(a, b)}); emap.emplace(ekey(a, b), g); }
            else g = it->second;
            cedge[nv * c + i] = g; cflip[nv * c + i] = a > b;
        }
    P.nedges = static_cast<int>(edges.size());
    P.raw_dim = np + P.nedges * de + nc * nint;
    P.cell_raw.assign(static_cast<std::size_t>(nc) * n, -1);
    for (int c = 0; c < nc; ++c)
    {
        int *d = &P.cell_raw[static_cast<std::size_t>(c) * n];
        for (int v = 0; v < nv; ++v) d[v] = M.tri[nv * c + v];
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < de; ++j)
                d[nv + i * de + j] = np + cedge[nv * c + i] * de + (cflip[nv * c + i] ? de - 1 - j : j);
        for (int j = 0; j < nint; ++j) d[f0 + j] = np + P.nedges * de + c * nint + j;
    }
    // node coordinates (space_2d.hpp:475-490)
    P.raw_xy.assign(2 * static_cast<std::size_t>(P.raw_dim), std::numeric_limits<double>::quiet_NaN());
    std::vector<char> done(static_cast<std::size_t>(P.raw_dim), 0);
    for (int v = 0; v < np; ++v) { P.raw_xy[2 * v] = M.xy[2 * v]; P.raw_xy[2 * v + 1] = M.xy[2 * v + 1]; done[v] = 1; }
    for (int c = 0; c < nc; ++c)
        for (int l = nv; l < n; ++l)
        {
            const int r = P.cell_raw[static_cast<std::size_t>(c) * n + l];
            if (done[r]) continue;
            if (nv == 4)
                quad_map(M, c, P.node_xi(l), P.node_eta(l), P.raw_xy[2 * r], P.raw_xy[2 * r + 1], nullptr);
            else
            {
                const double *J = &P.J[4 * c];
                const double ox = M.xy[2 * M.tri[3 * c]], oy = M.xy[2 * M.tri[3 * c] + 1];
                P.raw_xy[2 * r] = ox + J[0] * P.ref->nx[l] + J[1] * P.ref->ny[l];
                P.raw_xy[2 * r + 1] = oy + J[2] * P.ref->nx[l] + J[3] * P.ref->ny[l];
            }
            done[r] = 1;
        }
    // Dirichlet facets: exterior edges and segment edges, in cell order (space_2d.hpp:428-446, 491-498)
    std::vector<char> fixed(static_cast<std::size_t>(P.raw_dim), 0), seen(edges.size(), 0);
    for (int c = 0; c < nc; ++c)
        for (int i = 0; i < nv; ++i)
        {
            const int g = cedge[nv * c + i];
            int a, b;
            M.edge(c, i, a, b);
            const bool ext = M.nbr[nv * c + i] < 0;
            const int m = M.segment_marker(a, b);
            if (!ext && m < 0) continue;
            if (!ext && seen[g]) continue;
            seen[g] = 1;
            const int marker = m >= 0 ? m : 0;
            const bool on = (all_exterior && ext) || std::binary_search(dirichlet.begin(), dirichlet.end(), marker);
            if (!on) continue;
            fixed[edges[g][0]] = fixed[edges[g][1]] = 1;
            for (int j = 0; j < de; ++j) fixed[np + g * de + j] = 1;
        }
    // adapted numbering: raw order, the fixed nodes skipped (space_2d.hpp:506-526)
    P.raw2ad.assign(static_cast<std::size_t>(P.raw_dim), -1);
    P.ad2raw.clear();
    for (int r = 0; r < P.raw_dim; ++r)
        if (!fixed[r]) { P.raw2ad[r] = static_cast<int>(P.ad2raw.size()); P.ad2raw.push_back(r); }
    P.dim = static_cast<int>(P.ad2raw.size());
    P.cell_dof.resize(P.cell_raw.size());
    for (std::size_t i = 0; i < P.cell_raw.size(); ++i) P.cell_dof[i] = P.raw2ad[P.cell_raw[i]];
}

// -------------------------------------------------------------------------------- pattern
// sparse/csr.hpp:84-125 (from_cells: rows -> cells by counting, then per row sort + unique)
void pattern(const Problem &P, std::vector<int> &ip, std::vector<int> &ix)
{
    const int n = P.dim, nc = P.mesh.nt(), nl = P.nloc;
    std::vector<std::vector<int>> rows(static_cast<std::size_t>(n));
    for (int c = 0; c < nc; ++c)
        for (int a = 0; a < nl; ++a)
        {
            const int i = P.cell_dof[static_cast<std::size_t>(c) * nl + a];
            if (i < 0) continue;
            for (int b = 0; b < nl; ++b)
            {
                const int j = P.cell_dof[static_cast<std::size_t>(c) * nl + b];
                if (j >= 0) rows[i].push_back(j);
            }
        }
    ip.assign(static_cast<std::size_t>(n) + 1, 0);
    ix.clear();
    for (int i = 0; i < n; ++i)
    {
        std::sort(rows[i].begin(), rows[i].end());
        rows[i].erase(std::unique(rows[i].begin(), rows[i].end()), rows[i].end());
        ip[i + 1] = ip[i] + static_cast<int>(rows[i].size());
        ix.insert(ix.end(), rows[i].begin(), rows[i].end());
    }
}

// ------------------------------------------------------------------------------- integrand
// forms/integrand.hpp:33-46 (op codes), 300-342 (run), 159-163 (sech)
double run_program(int ninstr, const int *code, int nreg, const double *consts, double x, double y, std::vector<double> &reg,
                   double nx = 0.0, double ny = 0.0, const double *const *inputs = nullptr, std::size_t pt = 0)
{
    reg.assign(static_cast<std::size_t>(nreg), 0.0);
    int res = -1;
    for (int k = 0; k < ninstr; ++k)
    {
        const int op = code[4 * k], o = code[4 * k + 1], a = code[4 * k + 2], b = code[4 * k + 3];
        double v = 0.0;
        switch (op)
        {
        case 0: v = consts[a]; break;
        case 1: v = x; break;
        case 2: v = y; break;
        case 3: v = (a < 0 ? -1.0 : 1.0) * nx; break;     // integrand.hpp:314-315: the normal, a the side's sign
        case 4: v = (a < 0 ? -1.0 : 1.0) * ny; break;
        case 5:                                            // integrand.hpp:316: input a at the point
            if (!inputs) throw std::invalid_argument("integrand: the program reads an input");
            v = inputs[a][pt];
            break;
        case 6: v = reg[a] + reg[b]; break;
        case 7: v = reg[a] * reg[b]; break;
        case 8:
        {
            const double e = consts[b], r = reg[a];
            v = e == 2.0 ? r * r : e == -1.0 ? 1.0 / r : e == 0.5 ? std::sqrt(r) : std::pow(r, e);
            break;
        }
        case 9: v = std::sin(reg[a]); break;
        case 10: v = std::cos(reg[a]); break;
        case 11: v = std::exp(reg[a]); break;
        case 12: v = std::log(reg[a]); break;
        case 13: v = std::tanh(reg[a]); break;
        case 14: v = std::sqrt(reg[a]); break;
        case 15: v = std::sinh(reg[a]); break;
        case 16: v = std::cosh(reg[a]); break;
        case 17: { const double e = std::exp(-std::abs(reg[a])); v = 2.0 * e / (1.0 + e * e); break; }
        case 18: v = std::abs(reg[a]); break;
        case 19: v = reg[a] > 0.0 ? 1.0 : (reg[a] < 0.0 ? -1.0 : 0.0); break;
        default: throw std::invalid_argument("integrand: op not supported by the oracle");
        }
        reg[o] = v;
        res = o;
    }
    return res < 0 ? 0.0 : reg[res];
}

// ------------------------------------------------------------------------------- assembly
struct Quad {
    int nq = 0;
    std::vector<double> tab;              // nq x 3n reference tabulation
    std::vector<double> x, y, w;          // nc x nq (cache_2d.hpp:98-118)
    std::vector<double> Ji;               // nc x nq x 4 on quadrilaterals (J varies in the cell)
};

Quad make_quad(const Problem &P, int degree)
{
    Quad Q;
    std::vector<double> qx, qy, qw;
    if (P.quad()) quad_rule(degree, qx, qy, qw);
    else triangle_rule(degree, qx, qy, qw);
    Q.nq = static_cast<int>(qw.size());
    const int n = P.nloc, nc = P.mesh.nt();
    Q.tab.resize(static_cast<std::size_t>(Q.nq) * 3 * n);
    for (int q = 0; q < Q.nq; ++q) P.ref_eval(qx[q], qy[q], &Q.tab[static_cast<std::size_t>(q) * 3 * n]);
    Q.x.resize(static_cast<std::size_t>(nc) * Q.nq); Q.y.resize(Q.x.size()); Q.w.resize(Q.x.size());
    if (P.quad())
    {
        // cache_2d.hpp:98-118 with the bilinear map: J, det, w, J^{-1} per point
        Q.Ji.resize(4 * Q.x.size());
        for (int c = 0; c < nc; ++c)
            for (int q = 0; q < Q.nq; ++q)
            {
                const std::size_t k = static_cast<std::size_t>(c) * Q.nq + q;
                double J[4];
                quad_map(P.mesh, c, qx[q], qy[q], Q.x[k], Q.y[k], J);
                const double d = J[0] * J[3] - J[1] * J[2];
                Q.w[k] = qw[q] * std::abs(d);
                double *Ji = &Q.Ji[4 * k];
                Ji[0] = J[3] / d; Ji[1] = -J[1] / d;
                Ji[2] = -J[2] / d; Ji[3] = J[0] / d;
            }
        return Q;
    }
    for (int c = 0; c < nc; ++c)
    {
        const double *J = &P.J[4 * c];
        const double ox = P.mesh.xy[2 * P.mesh.tri[3 * c]], oy = P.mesh.xy[2 * P.mesh.tri[3 * c] + 1];
        const double d = J[0] * J[3] - J[1] * J[2];
        for (int q = 0; q < Q.nq; ++q)
        {
            const std::size_t k = static_cast<std::size_t>(c) * Q.nq + q;
            Q.x[k] = ox + J[0] * qx[q] + J[1] * qy[q];
            Q.y[k] = oy + J[2] * qx[q] + J[3] * qy[q];
            Q.w[k] = qw[q] * std::abs(d);
        }
    }
    return Q;
}

// cache_2d.hpp:211-226: code 0 value, 1 d/dx, 2 d/dy (J^{-1} recomputed per point from J, as there)
void phys_row(const Problem &P, const Quad &Q, int c, int q, int code, double *out)
{
    const int n = P.nloc;
    const double *R = &Q.tab[static_cast<std::size_t>(q) * 3 * n];
    if (code == 0) { std::copy(R, R + n, out); return; }
    double Ji[4];
    if (P.quad()) std::copy_n(&Q.Ji[4 * (static_cast<std::size_t>(c) * Q.nq + q)], 4, Ji);
    else
    {
        const double *J = &P.J[4 * c];
        const double d = J[0] * J[3] - J[1] * J[2];
        Ji[0] = J[3] / d; Ji[1] = -J[1] / d; Ji[2] = -J[2] / d; Ji[3] = J[0] / d;
    }
    const double a = code == 1 ? Ji[0] : Ji[1], b = code == 1 ? Ji[2] : Ji[3];
    for (int l = 0; l < n; ++l) out[l] = a * R[n + l] + b * R[2 * n + l];
}

// fe/cache_2d.hpp:245-288 (Cache2D::eval_multi): a known field with coefficients c (dim of them)
// at every point, out[i * np + e * nq + q] for codes[i] (0 value, 1 d/dx, 2 d/dy).  The cell's
// coefficients are gathered once (eliminated nodes read 0); the value alone is summed over the
// local functions, and with any derivative asked for the three sums run together, each local
// function's x and y rows formed from J^{-1} as the cache does.
void eval_field(const Problem &P, const Quad &Q, const double *c, int ncodes, const int *codes, double *out)
{
    bool deriv = false;
    for (int i = 0; i < ncodes; ++i)
    {
        if (codes[i] < 0 || codes[i] > 2) throw std::invalid_argument("derivative code is 0 (value), 1 (d/dx) or 2 (d/dy)");
        deriv = deriv || codes[i] > 0;
    }
    const int n = P.nloc, nc = P.mesh.nt();
    const std::size_t np = static_cast<std::size_t>(nc) * Q.nq;
    std::vector<double> loc(static_cast<std::size_t>(n));
    for (int e = 0; e < nc; ++e)
    {
        const int *d = &P.cell_dof[static_cast<std::size_t>(e) * n];
        for (int l = 0; l < n; ++l) loc[l] = d[l] >= 0 ? c[d[l]] : 0.0;
        for (int q = 0; q < Q.nq; ++q)
        {
            const double *R = &Q.tab[static_cast<std::size_t>(q) * 3 * n];
            const std::size_t k = static_cast<std::size_t>(e) * Q.nq + q;
            double s[3] = {0.0, 0.0, 0.0};
            if (!deriv)
                for (int l = 0; l < n; ++l) s[0] += R[l] * loc[l];
            else
            {
                double Ji[4];
                if (P.quad()) std::copy_n(&Q.Ji[4 * k], 4, Ji);
                else
                {
                    const double *J = &P.J[4 * e];
                    const double dt = J[0] * J[3] - J[1] * J[2];
                    Ji[0] = J[3] / dt; Ji[1] = -J[1] / dt; Ji[2] = -J[2] / dt; Ji[3] = J[0] / dt;
                }
                for (int l = 0; l < n; ++l)
                {
                    s[0] += R[l] * loc[l];
                    const double rx = Ji[0] * R[n + l] + Ji[2] * R[2 * n + l];
                    s[1] += rx * loc[l];
                    const double ry = Ji[1] * R[n + l] + Ji[3] * R[2 * n + l];
                    s[2] += ry * loc[l];
                }
            }
            for (int i = 0; i < ncodes; ++i) out[static_cast<std::size_t>(i) * np + k] = s[codes[i]];
        }
    }
}

// forms/assembly_2d.hpp:102-126 (element matrix: terms, points, rows, columns, zero products
// skipped) and 219-305 (the entries added to the CSR values in ascending element order).
void assemble_matrix(const Problem &P, const Quad &Q, int nt, const int *a, const int *b, const double *coef,
                     const std::vector<int> &ip, const std::vector<int> &ix, std::vector<double> &data, bool zero = true)
{
    const int n = P.nloc, nq = Q.nq, nc = P.mesh.nt();
    const std::size_t np = static_cast<std::size_t>(nc) * nq;
    if (zero) data.assign(static_cast<std::size_t>(ip.back()), 0.0);
    std::vector<double> Ke(static_cast<std::size_t>(n) * n), RT(static_cast<std::size_t>(3) * nq * n);
    for (int c = 0; c < nc; ++c)
    {
        for (int m = 0; m < 3; ++m)
            for (int q = 0; q < nq; ++q) phys_row(P, Q, c, q, m, &RT[(static_cast<std::size_t>(m) * nq + q) * n]);
        std::fill(Ke.begin(), Ke.end(), 0.0);
        for (int t = 0; t < nt; ++t)
            for (int q = 0; q < nq; ++q)
            {
                const std::size_t kk = static_cast<std::size_t>(c) * nq + q;
                const double s = Q.w[kk] * coef[t * np + kk];
                if (s == 0.0) continue;
                const double *ra = &RT[(static_cast<std::size_t>(a[t]) * nq + q) * n];
                const double *rb = &RT[(static_cast<std::size_t>(b[t]) * nq + q) * n];
                for (int i = 0; i < n; ++i)
                {
                    const double sa = s * ra[i];
                    if (sa == 0.0) continue;
                    for (int j = 0; j < n; ++j) Ke[i * n + j] += sa * rb[j];
                }
            }
        const int *d = &P.cell_dof[static_cast<std::size_t>(c) * n];
        for (int i = 0; i < n; ++i)
        {
            if (d[i] < 0) continue;
            for (int j = 0; j < n; ++j)
            {
                if (d[j] < 0) continue;
                const int *lo = ix.data() + ip[d[i]], *hi = ix.data() + ip[d[i] + 1];
                const int *p = std::lower_bound(lo, hi, d[j]);
                if (p == hi || *p != d[j]) throw std::logic_error("assembly: entry outside the pattern");
                data[p - ix.data()] += Ke[i * n + j];
            }
        }
    }
}

// forms/assembly_2d.hpp:173-211 and assembler_2d.hpp:166-188 (f = 0 + the group's block)
void assemble_vector(const Problem &P, const Quad &Q, int nt, const int *a, const double *coef, double *f, bool add = false)
{
    const int n = P.nloc, nq = Q.nq, nc = P.mesh.nt();
    const std::size_t np = static_cast<std::size_t>(nc) * nq;
    std::vector<double> block(static_cast<std::size_t>(P.dim), 0.0), fe(static_cast<std::size_t>(n)), r(static_cast<std::size_t>(n));
    for (int c = 0; c < nc; ++c)
    {
        std::fill(fe.begin(), fe.end(), 0.0);
        for (int t = 0; t < nt; ++t)
            for (int q = 0; q < nq; ++q)
            {
                const std::size_t kk = static_cast<std::size_t>(c) * nq + q;
                const double s = Q.w[kk] * coef[t * np + kk];
                if (s == 0.0) continue;
                phys_row(P, Q, c, q, a[t], r.data());
                for (int l = 0; l < n; ++l) fe[l] += s * r[l];
            }
        const int *d = &P.cell_dof[static_cast<std::size_t>(c) * n];
        for (int l = 0; l < n; ++l) if (d[l] >= 0) block[d[l]] += fe[l];
    }
    for (int i = 0; i < P.dim; ++i) f[i] = add ? f[i] + block[i] : 0.0 + block[i];
}

// --------------------------------------------------------------------- boundary facets (ds)
// fe/cache_2d.hpp:124-183: the exterior facets with a marker in `markers` (all of them with
// `all`), taken from the facet list of fe/space_2d.hpp:432-448 (cells ascending, local edges
// ascending); Gauss-Legendre points along each edge, max(1, (d + 2) / 2) of them on triangles,
// (d + 3) / 2 on quadrilaterals; per point the mapped point, 0.5 w_q |edge|, the outward normal
// (dy, -dx) / |edge| and J^{-1}; the basis tabulated once per local edge.
struct FacetQuad {
    int nq = 0, nent = 0;
    std::vector<int> cell, var;
    std::vector<double> tab;                      // (edge, point) x 3n
    std::vector<double> x, y, w, nx, ny, Ji;      // nent x nq (Ji x 4)
};

FacetQuad make_facet_quad(const Problem &P, int degree, std::vector<int> markers, bool all)
{
    std::sort(markers.begin(), markers.end());
    const Mesh &M = P.mesh;
    const int nv = M.nv, n = P.nloc;
    FacetQuad F;
    std::vector<double> g, gw;
    gauss_legendre(std::max(1, (degree + (nv == 4 ? 3 : 2)) / 2), g, gw);
    F.nq = static_cast<int>(g.size());
    const int nq = F.nq;
    // reference vertices (lagrange_space_2d.hpp:72-76, lagrange_space_quad.hpp:146-150) and the
    // local edges' end vertices (edge_vertices)
    const double tv[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
    const double qv[4][2] = {{0.0, 0.0}, {1.0, 0.0}, {1.0, 1.0}, {0.0, 1.0}};
    std::vector<double> fx(static_cast<std::size_t>(nv) * nq), fy(fx.size());
    F.tab.resize(static_cast<std::size_t>(nv) * nq * 3 * n);
    for (int i = 0; i < nv; ++i)
    {
        const int va = nv == 4 ? i : (i + 1) % 3, vb = nv == 4 ? (i + 1) % 4 : (i + 2) % 3;
        const double *A = nv == 4 ? qv[va] : tv[va], *B = nv == 4 ? qv[vb] : tv[vb];
        for (int q = 0; q < nq; ++q)
        {
            const double t = 0.5 * (g[q] + 1.0);
            const std::size_t k = static_cast<std::size_t>(i) * nq + q;
            fx[k] = (1.0 - t) * A[0] + t * B[0];
            fy[k] = (1.0 - t) * A[1] + t * B[1];
            P.ref_eval(fx[k], fy[k], &F.tab[k * 3 * n]);
        }
    }
    for (int c = 0; c < M.nt(); ++c)
        for (int i = 0; i < nv; ++i)
        {
            if (M.nbr[static_cast<std::size_t>(c) * nv + i] >= 0) continue;
            int a = 0, b = 0;
            M.edge(c, i, a, b);
            const int m = M.segment_marker(a, b);
            if (!all && !std::binary_search(markers.begin(), markers.end(), m >= 0 ? m : 0)) continue;
            F.cell.push_back(c);
            F.var.push_back(i);
        }
    F.nent = static_cast<int>(F.cell.size());
    const std::size_t np = static_cast<std::size_t>(F.nent) * nq;
    F.x.resize(np); F.y.resize(np); F.w.resize(np); F.nx.resize(np); F.ny.resize(np); F.Ji.resize(4 * np);
    for (int e = 0; e < F.nent; ++e)
    {
        const int c = F.cell[e], i = F.var[e];
        int a = 0, b = 0;
        M.edge(c, i, a, b);
        const double dx = M.xy[2 * b] - M.xy[2 * a], dy = M.xy[2 * b + 1] - M.xy[2 * a + 1];
        const double len = std::hypot(dx, dy);
        for (int q = 0; q < nq; ++q)
        {
            const std::size_t k = static_cast<std::size_t>(e) * nq + q;
            const double xi = fx[static_cast<std::size_t>(i) * nq + q], eta = fy[static_cast<std::size_t>(i) * nq + q];
            double J[4];
            if (P.quad()) quad_map(M, c, xi, eta, F.x[k], F.y[k], J);
            else
            {
                std::copy_n(&P.J[4 * static_cast<std::size_t>(c)], 4, J);
                const double ox = M.xy[2 * M.tri[3 * c]], oy = M.xy[2 * M.tri[3 * c] + 1];
                F.x[k] = ox + J[0] * xi + J[1] * eta;
                F.y[k] = oy + J[2] * xi + J[3] * eta;
            }
            F.w[k] = 0.5 * gw[q] * len;
            F.nx[k] = dy / len;
            F.ny[k] = -dx / len;
            const double d = J[0] * J[3] - J[1] * J[2];
            double *Ji = &F.Ji[4 * k];
            Ji[0] = J[3] / d; Ji[1] = -J[1] / d;
            Ji[2] = -J[2] / d; Ji[3] = J[0] / d;
        }
    }
    return F;
}

void facet_row(const Problem &P, const FacetQuad &F, int e, int q, int code, double *out)
{
    const int n = P.nloc;
    const double *R = &F.tab[(static_cast<std::size_t>(F.var[e]) * F.nq + q) * 3 * n];
    if (code == 0) { std::copy(R, R + n, out); return; }
    const double *Ji = &F.Ji[4 * (static_cast<std::size_t>(e) * F.nq + q)];
    const double a = code == 1 ? Ji[0] : Ji[1], b = code == 1 ? Ji[2] : Ji[3];
    for (int l = 0; l < n; ++l) out[l] = a * R[n + l] + b * R[2 * n + l];
}

// assemble_matrix_2d on a facet cache: each facet's element matrix (the loop order and zero skips
// of assembly_2d.hpp:102-126) added into data in ascending facet order (data is not zeroed: a
// form's groups add into one matrix, assembler_2d.hpp:249-266)
void assemble_matrix_facets(const Problem &P, const FacetQuad &F, int nt, const int *a, const int *b, const double *coef,
                            const std::vector<int> &ip, const std::vector<int> &ix, std::vector<double> &data)
{
    const int n = P.nloc, nq = F.nq;
    const std::size_t np = static_cast<std::size_t>(F.nent) * nq;
    std::vector<double> Ke(static_cast<std::size_t>(n) * n), RT(static_cast<std::size_t>(3) * nq * n);
    for (int e = 0; e < F.nent; ++e)
    {
        for (int m = 0; m < 3; ++m)
            for (int q = 0; q < nq; ++q) facet_row(P, F, e, q, m, &RT[(static_cast<std::size_t>(m) * nq + q) * n]);
        std::fill(Ke.begin(), Ke.end(), 0.0);
        for (int t = 0; t < nt; ++t)
            for (int q = 0; q < nq; ++q)
            {
                const std::size_t kk = static_cast<std::size_t>(e) * nq + q;
                const double s = F.w[kk] * coef[t * np + kk];
                if (s == 0.0) continue;
                const double *ra = &RT[(static_cast<std::size_t>(a[t]) * nq + q) * n];
                const double *rb = &RT[(static_cast<std::size_t>(b[t]) * nq + q) * n];
                for (int i = 0; i < n; ++i)
                {
                    const double sa = s * ra[i];
                    if (sa == 0.0) continue;
                    for (int j = 0; j < n; ++j) Ke[i * n + j] += sa * rb[j];
                }
            }
        const int *d = &P.cell_dof[static_cast<std::size_t>(F.cell[e]) * n];
        for (int i = 0; i < n; ++i)
        {
            if (d[i] < 0) continue;
            for (int j = 0; j < n; ++j)
            {
                if (d[j] < 0) continue;
                const int *lo = ix.data() + ip[d[i]], *hi = ix.data() + ip[d[i] + 1];
                const int *p = std::lower_bound(lo, hi, d[j]);
                if (p == hi || *p != d[j]) throw std::logic_error("assembly: entry outside the pattern");
                data[p - ix.data()] += Ke[i * n + j];
            }
        }
    }
}

// assemble_vector_2d on a facet cache into the group's own block, then f += block
// (assembler_2d.hpp:166-188)
void assemble_vector_facets(const Problem &P, const FacetQuad &F, int nt, const int *a, const double *coef, double *f)
{
    const int n = P.nloc, nq = F.nq;
    const std::size_t np = static_cast<std::size_t>(F.nent) * nq;
    std::vector<double> block(static_cast<std::size_t>(P.dim), 0.0), fe(static_cast<std::size_t>(n)), r(static_cast<std::size_t>(n));
    for (int e = 0; e < F.nent; ++e)
    {
        std::fill(fe.begin(), fe.end(), 0.0);
        for (int t = 0; t < nt; ++t)
            for (int q = 0; q < nq; ++q)
            {
                const std::size_t kk = static_cast<std::size_t>(e) * nq + q;
                const double s = F.w[kk] * coef[t * np + kk];
                if (s == 0.0) continue;
                facet_row(P, F, e, q, a[t], r.data());
                for (int l = 0; l < n; ++l) fe[l] += s * r[l];
            }
        const int *d = &P.cell_dof[static_cast<std::size_t>(F.cell[e]) * n];
        for (int l = 0; l < n; ++l) if (d[l] >= 0) block[d[l]] += fe[l];
    }
    for (int i = 0; i < P.dim; ++i) f[i] += block[i];
}

// ------------------------------------------------------------------------- vectors, sums
constexpr std::size_t CHUNK = 4096;       // util/omp.hpp:91

// util/omp.hpp:135-160: eight lanes s[i mod 8], the tail into s[0], fixed combination
double dot_chunk(const double *x, const double *y, std::size_t lo, std::size_t hi)
{
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::size_t i = lo;
    for (; i + 8 <= hi; i += 8)
        for (int k = 0; k < 8; ++k) s[k] += x[i + k] * y[i + k];
    for (; i < hi; ++i) s[0] += x[i] * y[i];
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// util/omp.hpp:176-192: chunk sums added left to right from zero
double ddot(const double *x, const double *y, std::size_t n)
{
    if (n <= CHUNK) return dot_chunk(x, y, 0, n);
    const long nch = static_cast<long>((n + CHUNK - 1) / CHUNK);
    std::vector<double> part(static_cast<std::size_t>(nch));
#pragma omp parallel for schedule(static) if (n > (1u << 20))
    for (long c = 0; c < nch; ++c)
        part[c] = dot_chunk(x, y, c * CHUNK, std::min(n, (c + 1) * CHUNK));
    double s = 0.0;
    for (double v : part) s += v;
    return s;
}

// sparse/csr.hpp:203-215: one row at a time, columns in stored order
void spmv(int n, const int *ip, const int *ix, const double *v, const double *x, double *y)
{
#pragma omp parallel for schedule(static) if (n > 4096)
    for (int i = 0; i < n; ++i)
    {
        double s = 0.0;
        for (int p = ip[i]; p < ip[i + 1]; ++p) s += v[p] * x[ix[p]];
        y[i] = s;
    }
}

// krylov/cg.hpp:33-102 and 174-208: x += alpha p, r -= alpha Ap, with ||r||^2 (and r^T D^{-1} r)
void cg_update(double *x, double *r, const double *p, const double *Ap, const double *inv, double alpha,
               std::size_t n, double &rr, double &rz)
{
    const long nch = static_cast<long>((n + CHUNK - 1) / CHUNK);
    std::vector<double> ps(static_cast<std::size_t>(std::max(nch, 1L))), pq(ps.size());
#pragma omp parallel for schedule(static) if (n > (1u << 20))
    for (long c = 0; c < std::max(nch, 1L); ++c)
    {
        const std::size_t lo = c * CHUNK, hi = std::min(n, (c + 1) * CHUNK);
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        std::size_t i = lo;
        for (; i + 8 <= hi; i += 8)
            for (int k = 0; k < 8; ++k)
            {
                r[i + k] = r[i + k] - alpha * Ap[i + k];
                x[i + k] = x[i + k] + alpha * p[i + k];
                const double e = r[i + k] * r[i + k];
                s[k] += e;
                if (inv) q[k] += e * inv[i + k];
            }
        for (; i < hi; ++i)
        {
            r[i] = r[i] - alpha * Ap[i];
            x[i] = x[i] + alpha * p[i];
            const double e = r[i] * r[i];
            s[0] += e;
            if (inv) q[0] += e * inv[i];
        }
        ps[c] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
        pq[c] = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
    }
    if (n <= CHUNK) { rr = ps[0]; rz = pq[0]; return; }
    rr = 0.0; rz = 0.0;
    for (long c = 0; c < nch; ++c) { rr += ps[c]; rz += pq[c]; }
}

struct CSR {
    int n = 0;
    std::vector<int> ip, ix;
    std::vector<double> v;
};

// sparse/amd.hpp:64-146 (symmetric graph without the diagonal, rows increasing) and 166-234
// (George-Liu pseudo-peripheral start, Cuthill-McKee by increasing degree, reversed)
std::vector<int> rcm(int n, const int *ip, const int *ix)
{
    std::vector<std::vector<int>> adj(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i)
        for (int p = ip[i]; p < ip[i + 1]; ++p)
        {
            const int j = ix[p];
            if (j < 0 || j >= n) throw std::invalid_argument("rcm: column out of range");
            if (j == i) continue;
            adj[i].push_back(j); adj[j].push_back(i);
        }
    for (auto &a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
    std::vector<int> deg(static_cast<std::size_t>(n)), level(static_cast<std::size_t>(n), -1), order, queue(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) deg[i] = static_cast<int>(adj[i].size());
    std::vector<char> done(static_cast<std::size_t>(n), 0);
    auto bfs = [&](int s, int &fl, int &end) {
        std::vector<int> touched{s};
        int head = 0, tail = 0, ecc = 0;
        queue[tail++] = s; level[s] = 0; fl = 0;
        while (head < tail)
        {
            const int v = queue[head++];
            if (level[v] > ecc) { ecc = level[v]; fl = head - 1; }
            for (int w : adj[v])
            {
                if (done[w] || level[w] >= 0) continue;
                level[w] = level[v] + 1; queue[tail++] = w; touched.push_back(w);
            }
        }
        end = tail;
        for (int v : touched) level[v] = -1;
        return ecc;
    };
    std::vector<int> byd(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) byd[i] = i;
    std::stable_sort(byd.begin(), byd.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int> nb;
    for (int seed : byd)
    {
        if (done[seed]) continue;
        int s = seed, fl, end;
        int ecc = bfs(s, fl, end);
        for (int it = 0; it < 32; ++it)
        {
            int best = queue[fl];
            for (int t = fl; t < end; ++t) if (deg[queue[t]] < deg[best]) best = queue[t];
            int fl2, end2;
            const int e2 = bfs(best, fl2, end2);
            if (e2 <= ecc) break;
            s = best; ecc = e2;
            bfs(s, fl, end);
        }
        std::size_t head = order.size();
        order.push_back(s); done[s] = 1;
        while (head < order.size())
        {
            const int v = order[head++];
            nb.clear();
            for (int w : adj[v]) if (!done[w]) { nb.push_back(w); done[w] = 1; }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// sparse/csr.hpp:286-313: B(k, l) = A(p[k], p[l]), rows sorted by new column
CSR permute(const CSR &A, const std::vector<int> &p)
{
    const int n = A.n;
    std::vector<int> pinv(static_cast<std::size_t>(n));
    for (int k = 0; k < n; ++k) pinv[p[k]] = k;
    CSR B;
    B.n = n;
    B.ip.assign(static_cast<std::size_t>(n) + 1, 0);
    std::vector<std::pair<int, double>> row;
    for (int k = 0; k < n; ++k)
    {
        row.clear();
        for (int q = A.ip[p[k]]; q < A.ip[p[k] + 1]; ++q) row.emplace_back(pinv[A.ix[q]], A.v[q]);
        std::sort(row.begin(), row.end());
        for (auto &e : row) { B.ix.push_back(e.first); B.v.push_back(e.second); }
        B.ip[k + 1] = static_cast<int>(B.ix.size());
    }
    return B;
}

std::vector<int> diag_positions(const CSR &A)
{
    std::vector<int> d(static_cast<std::size_t>(A.n));
    for (int i = 0; i < A.n; ++i)
    {
        const int *lo = A.ix.data() + A.ip[i], *hi = A.ix.data() + A.ip[i + 1];
        const int *p = std::lower_bound(lo, hi, i);
        d[i] = (p != hi && *p == i) ? static_cast<int>(p - A.ix.data()) : -1;
        if (d[i] < 0 || A.v[d[i]] == 0.0) throw std::invalid_argument("zero diagonal entry in row " + std::to_string(i));
    }
    return d;
}

// sparse/precond.hpp: Jacobi (82-102), SSOR (105-148), ILU(0) (151-205), permuted apply (56-63)
struct Precond {
    int kind = 0;                     // 1 Jacobi, 2 ILU(0), 3 SSOR
    int n = 0;
    double omega = 1.0, min_pivot = 0.0;
    std::vector<int> perm, diag;
    std::vector<double> inv;
    CSR LU;
    mutable std::vector<double> rp, zp;

    Precond(const CSR &A, int kind_, int rcm_order, double om) : kind(kind_), n(A.n), omega(om)
    {
        if (kind == 1)
        {
            diag = diag_positions(A);
            inv.resize(static_cast<std::size_t>(n));
            for (int i = 0; i < n; ++i) inv[i] = 1.0 / A.v[diag[i]];
            return;
        }
        if (rcm_order) perm = rcm(n, A.ip.data(), A.ix.data());
        LU = perm.empty() ? A : permute(A, perm);
        diag = diag_positions(LU);
        if (kind == 3) return;
        std::vector<int> pos(static_cast<std::size_t>(n), -1);
        min_pivot = std::numeric_limits<double>::infinity();
        double *v = LU.v.data();
        for (int i = 0; i < n; ++i)
        {
            for (int p = LU.ip[i]; p < LU.ip[i + 1]; ++p) pos[LU.ix[p]] = p;
            for (int p = LU.ip[i]; p < diag[i]; ++p)
            {
                const int k = LU.ix[p];
                v[p] /= v[diag[k]];
                const double lik = v[p];
                for (int q = diag[k] + 1; q < LU.ip[k + 1]; ++q)
                {
                    const int t = pos[LU.ix[q]];
                    if (t >= 0) v[t] -= lik * v[q];
                }
            }
            for (int p = LU.ip[i]; p < LU.ip[i + 1]; ++p) pos[LU.ix[p]] = -1;
            const double d = v[diag[i]];
            if (d == 0.0) throw std::runtime_error("ILU(0): zero pivot in row " + std::to_string(i));
            min_pivot = std::min(min_pivot, d);
        }
    }
    void core(const double *r, double *z) const
    {
        const int *ip = LU.ip.data(), *ix = LU.ix.data();
        const double *v = LU.v.data();
        if (kind == 2)
        {
            for (int i = 0; i < n; ++i)
            {
                double s = r[i];
                for (int p = ip[i]; p < diag[i]; ++p) s -= v[p] * z[ix[p]];
                z[i] = s;
            }
            for (int i = n - 1; i >= 0; --i)
            {
                double s = z[i];
                for (int p = diag[i] + 1; p < ip[i + 1]; ++p) s -= v[p] * z[ix[p]];
                z[i] = s / v[diag[i]];
            }
            return;
        }
        const double w = omega;
        for (int i = 0; i < n; ++i)
        {
            double s = r[i];
            for (int p = ip[i]; p < diag[i]; ++p) s -= v[p] * z[ix[p]];
            z[i] = s * w / v[diag[i]];
        }
        const double c = (2.0 - w) / w;
        for (int i = 0; i < n; ++i) z[i] *= c * v[diag[i]] / w;
        for (int i = n - 1; i >= 0; --i)
        {
            double s = z[i];
            for (int p = diag[i] + 1; p < ip[i + 1]; ++p) s -= v[p] * z[ix[p]];
            z[i] = s * w / v[diag[i]];
        }
    }
    void apply(const double *r, double *z) const
    {
        if (kind == 1) { for (int i = 0; i < n; ++i) z[i] = inv[i] * r[i]; return; }
        if (perm.empty()) { core(r, z); return; }
        rp.resize(static_cast<std::size_t>(n)); zp.resize(static_cast<std::size_t>(n));
        for (int k = 0; k < n; ++k) rp[k] = r[perm[k]];
        core(rp.data(), zp.data());
        for (int k = 0; k < n; ++k) z[perm[k]] = zp[k];
    }
};

struct CGOut { bool converged = false, breakdown = false; int iterations = 0; double residual = 0.0; };

// krylov/cg.hpp:222-237 (final true residual and verdict)
void cg_finish(const CSR &A, std::vector<double> &x, const std::vector<double> &b, std::vector<double> &Ap,
               std::vector<double> &r, double nb, int it, int max_iter, double tol, CGOut &out)
{
    const std::size_t n = b.size();
    spmv(A.n, A.ip.data(), A.ix.data(), A.v.data(), x.data(), Ap.data());
    for (std::size_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    out.iterations = it;
    out.residual = std::sqrt(ddot(r.data(), r.data(), n)) / nb;
    out.converged = !out.breakdown && out.residual <= tol * (1.0 + 1e-6) + 1e-15;
    if (!out.converged && !out.breakdown && out.residual <= 10.0 * tol && it < max_iter) out.converged = true;
}

// krylov/cg.hpp:295-350 (cg_jacobi: inv null = no preconditioner)
CGOut cg_jacobi(const CSR &A, const double *inv, std::vector<double> &x, const std::vector<double> &b, int max_iter, double tol)
{
    const std::size_t n = b.size();
    CGOut out;
    const double nb = std::sqrt(ddot(b.data(), b.data(), n));
    if (nb == 0.0) { std::fill(x.begin(), x.end(), 0.0); out.converged = true; return out; }
    std::vector<double> r(n), p(n), Ap(n);
    spmv(A.n, A.ip.data(), A.ix.data(), A.v.data(), x.data(), Ap.data());
    for (std::size_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    double rn = std::sqrt(ddot(r.data(), r.data(), n));
    if (rn <= tol * nb) { out.converged = true; out.residual = rn / nb; return out; }
    double rz;
    if (inv) { for (std::size_t i = 0; i < n; ++i) p[i] = inv[i] * r[i]; rz = ddot(r.data(), p.data(), n); }
    else { p = r; rz = rn * rn; }
    int it = 0;
    while (it < max_iter)
    {
        if (!(rz > 0.0)) { out.breakdown = true; break; }
        spmv(A.n, A.ip.data(), A.ix.data(), A.v.data(), p.data(), Ap.data());
        const double pAp = ddot(p.data(), Ap.data(), n);
        if (!(pAp > 0.0)) { out.breakdown = true; break; }
        const double alpha = rz / pAp;
        double rr, rzn;
        cg_update(x.data(), r.data(), p.data(), Ap.data(), inv, alpha, n, rr, rzn);
        if (!inv) rzn = rr;
        rn = std::sqrt(rr);
        ++it;
        if (rn <= tol * nb) break;
        const double beta = rzn / rz;
        rz = rzn;
        if (inv) for (std::size_t i = 0; i < n; ++i) p[i] = inv[i] * r[i] + beta * p[i];
        else for (std::size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    }
    cg_finish(A, x, b, Ap, r, nb, it, max_iter, tol, out);
    return out;
}

// krylov/cg.hpp:245-287 (cg with any preconditioner)
CGOut cg_general(const CSR &A, const Precond &M, std::vector<double> &x, const std::vector<double> &b, int max_iter, double tol)
{
    const std::size_t n = b.size();
    CGOut out;
    const double nb = std::sqrt(ddot(b.data(), b.data(), n));
    if (nb == 0.0) { std::fill(x.begin(), x.end(), 0.0); out.converged = true; return out; }
    std::vector<double> r(n), z(n), p, Ap(n);
    spmv(A.n, A.ip.data(), A.ix.data(), A.v.data(), x.data(), Ap.data());
    for (std::size_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    double rn = std::sqrt(ddot(r.data(), r.data(), n));
    if (rn <= tol * nb) { out.converged = true; out.residual = rn / nb; return out; }
    M.apply(r.data(), z.data());
    p = z;
    double rz = ddot(r.data(), z.data(), n);
    int it = 0;
    while (it < max_iter)
    {
        if (!(rz > 0.0)) { out.breakdown = true; break; }
        spmv(A.n, A.ip.data(), A.ix.data(), A.v.data(), p.data(), Ap.data());
        const double pAp = ddot(p.data(), Ap.data(), n);
        if (!(pAp > 0.0)) { out.breakdown = true; break; }
        const double alpha = rz / pAp;
        double rr, unused;
        cg_update(x.data(), r.data(), p.data(), Ap.data(), nullptr, alpha, n, rr, unused);
        rn = std::sqrt(rr);
        ++it;
        if (rn <= tol * nb) break;
        M.apply(r.data(), z.data());
        const double rzn = ddot(r.data(), z.data(), n);
        const double beta = rzn / rz;
        rz = rzn;
        for (std::size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    cg_finish(A, x, b, Ap, r, nb, it, max_iter, tol, out);
    return out;
}

// ----------------------------------------------------------------------- GMRES / LGMRES
// vendor/csnewton/csnewton.hpp:87-156 (serial dot, norm, axpy, Givens) and 174-260 (gmres),
// 298-474 (lgmres); krylov/adapters.hpp:116-147 (tol rescaling, residuals recomputed).
double sdot(const std::vector<double> &u, const std::vector<double> &v)
{
    double s = 0.0;
    for (std::size_t i = 0; i < u.size(); ++i) s += u[i] * v[i];
    return s;
}
double snorm(const std::vector<double> &v) { return std::sqrt(sdot(v, v)); }
void gen_givens(double dx, double dy, double &cs, double &sn)
{
    const double ax = std::abs(dx), ay = std::abs(dy);
    if (ay == 0.0) { cs = 1.0; sn = 0.0; }
    else if (ax == 0.0) { cs = 0.0; sn = 1.0; }
    else { const double tau = std::sqrt(ax * ax + ay * ay); cs = ax / tau; sn = (dx / ax) * (ay / tau); }
}
void app_givens(double &dx, double &dy, double cs, double sn)
{
    const double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

struct GOut { bool converged = false; int iterations = 0; double residual = 0.0; };

template <class Op, class Pre>
GOut gmres(Op &&A, Pre &&M, std::vector<double> &x, const std::vector<double> &b, int restart, int max_iter, double tol)
{
    const int n = static_cast<int>(b.size()), ldh = restart + 1;
    if (restart <= 0 || max_iter <= 0) throw std::invalid_argument("gmres: restart and max_iter must be positive");
    std::vector<double> H(static_cast<std::size_t>(ldh) * restart, 0.0), cs(restart), sn(restart), s(restart + 1);
    std::vector<std::vector<double>> V(restart + 1, std::vector<double>(n));
    std::vector<double> w(n), r(n), z(n);
    auto update = [&](int k) {
        std::vector<double> y(s.begin(), s.begin() + k + 1);
        for (int i = k; i >= 0; --i)
        {
            y[i] /= H[i + i * ldh];
            for (int j = 0; j < i; ++j) y[j] -= H[j + i * ldh] * y[i];
        }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= k; ++j) x[i] += y[j] * V[j][i];
    };
    A(x, w);
    for (int i = 0; i < n; ++i) r[i] = b[i] - w[i];
    M(r, z);
    double normb = snorm(b);
    if (normb == 0.0) normb = 1.0;
    double beta = snorm(z);
    if (beta / normb < tol) return {true, 0, beta / normb};
    int total = 0;
    while (total < max_iter)
    {
        std::fill(H.begin(), H.end(), 0.0); std::fill(s.begin(), s.end(), 0.0);
        s[0] = beta;
        V[0] = z;
        for (double &e : V[0]) e *= 1.0 / beta;
        int i = 0;
        for (; i < restart && total < max_iter; ++i, ++total)
        {
            A(V[i], r);
            M(r, w);
            for (int k = 0; k <= i; ++k)
            {
                H[k + i * ldh] = sdot(w, V[k]);
                const double h = -H[k + i * ldh];
                for (int t = 0; t < n; ++t) w[t] += h * V[k][t];
            }
            const double hn = snorm(w);
            H[(i + 1) + i * ldh] = hn;
            if (hn == 0.0) { update(i); return {true, total + 1, 0.0}; }
            V[i + 1] = w;
            for (double &e : V[i + 1]) e *= 1.0 / hn;
            for (int k = 0; k < i; ++k) app_givens(H[k + i * ldh], H[(k + 1) + i * ldh], cs[k], sn[k]);
            gen_givens(H[i + i * ldh], H[(i + 1) + i * ldh], cs[i], sn[i]);
            app_givens(H[i + i * ldh], H[(i + 1) + i * ldh], cs[i], sn[i]);
            app_givens(s[i], s[i + 1], cs[i], sn[i]);
            const double est = std::abs(s[i + 1]) / normb;
            if (est < tol) { update(i); return {true, total + 1, est}; }
        }
        update(i - 1);
        A(x, w);
        for (int k = 0; k < n; ++k) r[k] = b[k] - w[k];
        M(r, z);
        beta = snorm(z);
        if (beta / normb < tol) return {true, total, beta / normb};
    }
    return {false, total, beta / normb};
}

template <class Op, class Pre>
GOut lgmres(Op &&A, Pre &&M, std::vector<double> &x, const std::vector<double> &b, int inner, int k_aug, int max_iter, double tol)
{
    const int n = static_cast<int>(b.size());
    if (inner <= 0 || max_iter <= 0) throw std::invalid_argument("lgmres: inner and max_iter must be positive");
    if (k_aug < 0) throw std::invalid_argument("lgmres: k_aug must be non-negative");
    std::vector<std::vector<double>> Zaug, MAZ;
    double normb = snorm(b);
    if (normb == 0.0) normb = 1.0;
    std::vector<double> Ax(n), r(n), q(n), w(n);
    int total = 0;
    auto arnoldi_col = [&](int j, int ldh, std::vector<double> &H, std::vector<std::vector<double>> &V,
                           std::vector<double> &cs, std::vector<double> &sn, std::vector<double> &s) {
        for (int i = 0; i <= j; ++i)
        {
            H[i + j * ldh] = sdot(w, V[i]);
            const double h = -H[i + j * ldh];
            for (int t = 0; t < n; ++t) w[t] += h * V[i][t];
        }
        const double hn = snorm(w);
        H[(j + 1) + j * ldh] = hn;
        if (hn > 0.0) { V[j + 1] = w; for (double &e : V[j + 1]) e *= 1.0 / hn; }
        for (int i = 0; i < j; ++i) app_givens(H[i + j * ldh], H[(i + 1) + j * ldh], cs[i], sn[i]);
        gen_givens(H[j + j * ldh], H[(j + 1) + j * ldh], cs[j], sn[j]);
        app_givens(H[j + j * ldh], H[(j + 1) + j * ldh], cs[j], sn[j]);
        app_givens(s[j], s[j + 1], cs[j], sn[j]);
        return hn;
    };
    while (total < max_iter)
    {
        const int k = static_cast<int>(Zaug.size()), m = inner + k, ldh = m + 1;
        A(x, Ax);
        for (int i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
        M(r, q);
        const double beta = snorm(q);
        if (beta / normb < tol) return {true, total, beta / normb};
        std::vector<double> H(static_cast<std::size_t>(ldh) * m, 0.0), cs(m), sn(m), s(m + 1, 0.0);
        std::vector<std::vector<double>> V(m + 1, std::vector<double>(n));
        V[0] = q;
        for (double &e : V[0]) e *= 1.0 / beta;
        s[0] = beta;
        int final_j = m - 1, built = 0;
        bool conv = false;
        const int inner_lim = std::min(inner, max_iter - total);
        for (int j = 0; j < inner_lim; ++j, ++total)
        {
            A(V[j], r);
            M(r, w);
            const double hn = arnoldi_col(j, ldh, H, V, cs, sn, s);
            built = j + 1;
            if (std::abs(s[j + 1]) / normb < tol || hn == 0.0) { final_j = j; conv = true; ++total; break; }
        }
        if (!conv)
        {
            const int aug_lim = std::min(k, max_iter - total);
            for (int aug = 0; aug < aug_lim; ++aug, ++total)
            {
                const int j = inner + aug;
                w = MAZ[aug];
                const double hn = arnoldi_col(j, ldh, H, V, cs, sn, s);
                built = j + 1;
                if (std::abs(s[j + 1]) / normb < tol || hn == 0.0) { final_j = j; conv = true; ++total; break; }
            }
        }
        if (!conv) final_j = built - 1;
        std::vector<double> y(s.begin(), s.begin() + final_j + 1);
        for (int i = final_j; i >= 0; --i)
        {
            y[i] /= H[i + i * ldh];
            for (int jj = 0; jj < i; ++jj) y[jj] -= H[jj + i * ldh] * y[i];
        }
        std::vector<double> znew(n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= final_j; ++j)
            {
                const double c = y[j] * (j < inner ? V[j][i] : Zaug[j - inner][i]);
                x[i] += c;
                znew[i] += c;
            }
        if (k_aug > 0)
        {
            std::vector<double> Az(n), MAz(n);
            A(znew, Az);
            M(Az, MAz);
            Zaug.insert(Zaug.begin(), znew);
            MAZ.insert(MAZ.begin(), MAz);
            if (static_cast<int>(Zaug.size()) > k_aug) { Zaug.pop_back(); MAZ.pop_back(); }
        }
        if (conv)
        {
            A(x, Ax);
            for (int i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
            M(r, q);
            return {true, total, snorm(q) / normb};
        }
    }
    A(x, Ax);
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
    M(r, q);
    return {false, total, snorm(q) / normb};
}

CSR make_csr(int n, const int *ip, const int *ix, const double *v)
{
    CSR A;
    A.n = n;
    A.ip.assign(ip, ip + n + 1);
    A.ix.assign(ix, ix + ip[n]);
    A.v.assign(v, v + ip[n]);
    return A;
}

} // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/// A structured rectangle mesh and a P_k space on it (diag: 0 right, 1 left, 2 alternate, 3 crossed).
void *orc_problem_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, int diag, int k,
                            int ndir, const int *dir, int all_exterior, int equispaced)
{
    Problem *out = nullptr;
    guard([&] {
        auto P = std::make_unique<Problem>();
        const int mk[4] = {1, 2, 3, 4};
        P->mesh = rectangle(x0, y0, x1, y1, nx, ny, diag, mk);
        build_space(*P, k, std::vector<int>(dir, dir + ndir), all_exterior != 0, equispaced != 0);
        out = P.release();
    });
    return out;
}

/// A mesh from arrays (structured.hpp:99-221) and a P_k space on it.
void *orc_problem_arrays(int np, const double *pts, int nt, const int *tris, int ns, const int *segs, const int *smarks,
                         int k, int ndir, const int *dir, int all_exterior, int equispaced)
{
    Problem *out = nullptr;
    guard([&] {
        auto P = std::make_unique<Problem>();
        P->mesh = build_mesh(std::vector<double>(pts, pts + 2 * np), std::vector<int>(tris, tris + 3 * nt),
                             std::vector<int>(segs, segs + 2 * ns), smarks ? std::vector<int>(smarks, smarks + ns) : std::vector<int>{}, 1);
        build_space(*P, k, std::vector<int>(dir, dir + ndir), all_exterior != 0, equispaced != 0);
        out = P.release();
    });
    return out;
}

/// A structured quadrilateral mesh of the rectangle (quadmesh.py:161-170) and a Q_k space on it.
void *orc_problem_quad_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, int k, int ndir, const int *dir,
                                 int all_exterior, int equispaced)
{
    Problem *out = nullptr;
    guard([&] {
        auto P = std::make_unique<Problem>();
        const int mk[4] = {1, 2, 3, 4};
        P->mesh = quad_rectangle(x0, y0, x1, y1, nx, ny, mk);
        build_space(*P, k, std::vector<int>(dir, dir + ndir), all_exterior != 0, equispaced != 0);
        out = P.release();
    });
    return out;
}

/// A quadrilateral mesh from arrays (quad_mesh.hpp:50-103) and a Q_k space on it.
void *orc_problem_quad_arrays(int np, const double *pts, int nq, const int *quads, int ns, const int *segs, const int *smarks,
                              int k, int ndir, const int *dir, int all_exterior, int equispaced)
{
    Problem *out = nullptr;
    guard([&] {
        auto P = std::make_unique<Problem>();
        P->mesh = build_quad_mesh(std::vector<double>(pts, pts + 2 * np), std::vector<int>(quads, quads + 4 * nq),
                                  std::vector<int>(segs, segs + 2 * ns), smarks ? std::vector<int>(smarks, smarks + ns) : std::vector<int>{},
                                  1);
        build_space(*P, k, std::vector<int>(dir, dir + ndir), all_exterior != 0, equispaced != 0);
        out = P.release();
    });
    return out;
}

void orc_problem_free(void *h) { delete static_cast<Problem *>(h); }

/// out: ncells, nloc, raw_dim, dim, nedges, npoints, nsegments
int orc_problem_sizes(void *h, int *out)
{
    return guard([&] {
        const Problem &P = *static_cast<Problem *>(h);
        out[0] = P.mesh.nt(); out[1] = P.nloc; out[2] = P.raw_dim; out[3] = P.dim; out[4] = P.nedges; out[5] = P.mesh.np();
        out[6] = static_cast<int>(P.mesh.smark.size());
    });
}

int orc_mesh_arrays(void *h, double *pts, int *tris, int *nbr, int *segs, int *smarks)
{
    return guard([&] {
        const Mesh &M = static_cast<Problem *>(h)->mesh;
        std::copy(M.xy.begin(), M.xy.end(), pts);
        std::copy(M.tri.begin(), M.tri.end(), tris);
        std::copy(M.nbr.begin(), M.nbr.end(), nbr);
        std::copy(M.seg.begin(), M.seg.end(), segs);
        std::copy(M.smark.begin(), M.smark.end(), smarks);
    });
}

int orc_numbering(void *h, int *cell_dof, int *cell_raw, int *ad2raw, double *raw_xy)
{
    return guard([&] {
        const Problem &P = *static_cast<Problem *>(h);
        std::copy(P.cell_dof.begin(), P.cell_dof.end(), cell_dof);
        std::copy(P.cell_raw.begin(), P.cell_raw.end(), cell_raw);
        std::copy(P.ad2raw.begin(), P.ad2raw.end(), ad2raw);
        std::copy(P.raw_xy.begin(), P.raw_xy.end(), raw_xy);
    });
}

long orc_pattern_nnz(void *h)
{
    long n = -1;
    guard([&] { std::vector<int> ip, ix; pattern(*static_cast<Problem *>(h), ip, ix); n = ip.back(); });
    return n;
}

int orc_pattern(void *h, int *indptr, int *indices)
{
    return guard([&] {
        std::vector<int> ip, ix;
        pattern(*static_cast<Problem *>(h), ip, ix);
        std::copy(ip.begin(), ip.end(), indptr);
        std::copy(ix.begin(), ix.end(), indices);
    });
}

int orc_quadrature(int degree, int *nq, double *xi, double *eta, double *w)
{
    return guard([&] {
        std::vector<double> a, b, c;
        triangle_rule(degree, a, b, c);
        *nq = static_cast<int>(c.size());
        std::copy(a.begin(), a.end(), xi); std::copy(b.begin(), b.end(), eta); std::copy(c.begin(), c.end(), w);
    });
}

int orc_basis_eval(int k, int equispaced, double xi, double eta, double *out)
{
    return guard([&] { RefElement R(k, equispaced != 0); R.eval(xi, eta, out); });
}

/// The quadrilateral rule of `degree` (tensor Gauss-Legendre on [0,1]^2).
int orc_quadrature_quad(int degree, int *nq, double *xi, double *eta, double *w)
{
    return guard([&] {
        std::vector<double> a, b, c;
        quad_rule(degree, a, b, c);
        *nq = static_cast<int>(c.size());
        std::copy(a.begin(), a.end(), xi); std::copy(b.begin(), b.end(), eta); std::copy(c.begin(), c.end(), w);
    });
}

/// Values, d/dxi, d/deta of the Q_k basis at (xi, eta): 3 (k+1)^2 numbers.
int orc_basis_eval_quad(int k, int equispaced, double xi, double eta, double *out)
{
    return guard([&] { RefQuad R(k, equispaced != 0); R.eval(xi, eta, out); });
}

/// The local nodes (xi, eta) of the P_k (quad = 0) or Q_k (quad = 1) reference element.
int orc_ref_nodes(int k, int equispaced, int quad, double *xi, double *eta)
{
    return guard([&] {
        if (quad)
        {
            RefQuad R(k, equispaced != 0);
            std::copy(R.nx.begin(), R.nx.end(), xi); std::copy(R.ny.begin(), R.ny.end(), eta);
        }
        else
        {
            RefElement R(k, equispaced != 0);
            std::copy(R.nx.begin(), R.nx.end(), xi); std::copy(R.ny.begin(), R.ny.end(), eta);
        }
    });
}

/// The p+1 Gauss-Lobatto points on [-1, 1].
int orc_lobatto_points(int p, double *out)
{
    return guard([&] {
        if (p < 1) throw std::invalid_argument("Gauss-Lobatto points need p >= 1");
        const std::vector<double> x = lobatto_points(p);
        std::copy(x.begin(), x.end(), out);
    });
}

int orc_cache_points(void *h, int degree, double *x, double *y, double *w)
{
    return guard([&] {
        Quad Q = make_quad(*static_cast<Problem *>(h), degree);
        std::copy(Q.x.begin(), Q.x.end(), x); std::copy(Q.y.begin(), Q.y.end(), y); std::copy(Q.w.begin(), Q.w.end(), w);
    });
}

int orc_integrand_eval(int ninstr, const int *code, int nreg, const double *consts, long npts, const double *xs,
                       const double *ys, double *out)
{
    return guard([&] {
        std::vector<double> reg;
        for (long i = 0; i < npts; ++i) out[i] = run_program(ninstr, code, nreg, consts, xs ? xs[i] : 0.0, ys ? ys[i] : 0.0, reg);
    });
}

/// The integrand with inputs: ninput arrays of npts values each, laid out one after another.
int orc_integrand_eval_in(int ninstr, const int *code, int nreg, const double *consts, long npts, const double *xs,
                          const double *ys, int ninput, const double *inputs, double *out)
{
    return guard([&] {
        std::vector<const double *> in(static_cast<std::size_t>(std::max(ninput, 0)));
        for (int i = 0; i < ninput; ++i) in[static_cast<std::size_t>(i)] = inputs + static_cast<std::size_t>(i) * npts;
        std::vector<double> reg;
        for (long i = 0; i < npts; ++i)
            out[i] = run_program(ninstr, code, nreg, consts, xs ? xs[i] : 0.0, ys ? ys[i] : 0.0, reg, 0.0, 0.0, in.data(),
                                 static_cast<std::size_t>(i));
    });
}

/// A known field at the cell points of the rule of a degree (Cache2D::eval_multi).
int orc_eval_field(void *h, int degree, const double *c, int ncodes, const int *codes, double *out)
{
    return guard([&] {
        const Problem &P = *static_cast<Problem *>(h);
        eval_field(P, make_quad(P, degree), c, ncodes, codes, out);
    });
}

int orc_assemble_matrix(void *h, int degree, int nterms, const int *a, const int *b, const double *coeffs, double *data)
{
    return guard([&] {
        const Problem &P = *static_cast<Problem *>(h);
        Quad Q = make_quad(P, degree);
        std::vector<int> ip, ix;
        pattern(P, ip, ix);
        std::vector<double> v;
        assemble_matrix(P, Q, nterms, a, b, coeffs, ip, ix, v);
        std::copy(v.begin(), v.end(), data);
    });
}

int orc_assemble_vector(void *h, int degree, int nterms, const int *a, const double *coeffs, double *f)
{
    return guard([&] {
        const Problem &P = *static_cast<Problem *>(h);
        Quad Q = make_quad(P, degree);
        assemble_vector(P, Q, nterms, a, coeffs, f);
    });
}

/// The point set of ds(markers) (every exterior facet with all): out = (facets, points per facet).
int orc_facet_sizes(void *h, int degree, int nmark, const int *marks, int all, int *out)
{
    return guard([&] {
        FacetQuad F = make_facet_quad(*static_cast<Problem *>(h), degree, std::vector<int>(marks, marks + nmark), all != 0);
        out[0] = F.nent; out[1] = F.nq;
    });
}

int orc_facet_points(void *h, int degree, int nmark, const int *marks, int all, double *x, double *y, double *w, double *nx,
                     double *ny)
{
    return guard([&] {
        FacetQuad F = make_facet_quad(*static_cast<Problem *>(h), degree, std::vector<int>(marks, marks + nmark), all != 0);
        std::copy(F.x.begin(), F.x.end(), x); std::copy(F.y.begin(), F.y.end(), y); std::copy(F.w.begin(), F.w.end(), w);
        std::copy(F.nx.begin(), F.nx.end(), nx); std::copy(F.ny.begin(), F.ny.end(), ny);
    });
}

/// A form's groups in order (forms/assembler_2d.hpp:166-188, 249-266): group g is dx (kind 0) or
/// ds (kind 1, markers marks[...] of count nmark[g], or every exterior facet with all[g]) at
/// degree[g], with nterms[g] terms (codes a, b) whose coefficient arrays (entries x points of the
/// group's point set) follow one another over all groups.  K is zeroed, then every group adds.
int orc_assemble_matrix_groups(void *h, int ngroups, const int *kind, const int *degree, const int *nmark, const int *marks,
                               const int *all, const int *nterms, const int *a, const int *b, const double *coeffs, double *data)
{
    return guard([&] {
        const Problem &P = *static_cast<Problem *>(h);
        std::vector<int> ip, ix;
        pattern(P, ip, ix);
        std::vector<double> v(static_cast<std::size_t>(ip.back()), 0.0);
        std::size_t mo = 0, to = 0, co = 0;
        for (int g = 0; g < ngroups; ++g)
        {
            if (kind[g] == 0)
            {
                Quad Q = make_quad(P, degree[g]);
                assemble_matrix(P, Q, nterms[g], a + to, b + to, coeffs + co, ip, ix, v, false);
                co += static_cast<std::size_t>(nterms[g]) * P.mesh.nt() * Q.nq;
            }
            else
            {
                FacetQuad F = make_facet_quad(P, degree[g], std::vector<int>(marks + mo, marks + mo + nmark[g]), all[g] != 0);
                assemble_matrix_facets(P, F, nterms[g], a + to, b + to, coeffs + co, ip, ix, v);
                co += static_cast<std::size_t>(nterms[g]) * F.nent * F.nq;
            }
            mo += nmark[g];
            to += nterms[g];
        }
        std::copy(v.begin(), v.end(), data);
    });
}

int orc_assemble_vector_groups(void *h, int ngroups, const int *kind, const int *degree, const int *nmark, const int *marks,
                               const int *all, const int *nterms, const int *a, const double *coeffs, double *f)
{
    return guard([&] {
        const Problem &P = *static_cast<Problem *>(h);
        std::fill(f, f + P.dim, 0.0);
        std::size_t mo = 0, to = 0, co = 0;
        for (int g = 0; g < ngroups; ++g)
        {
            if (kind[g] == 0)
            {
                Quad Q = make_quad(P, degree[g]);
                assemble_vector(P, Q, nterms[g], a + to, coeffs + co, f, true);
                co += static_cast<std::size_t>(nterms[g]) * P.mesh.nt() * Q.nq;
            }
            else
            {
                FacetQuad F = make_facet_quad(P, degree[g], std::vector<int>(marks + mo, marks + mo + nmark[g]), all[g] != 0);
                assemble_vector_facets(P, F, nterms[g], a + to, coeffs + co, f);
                co += static_cast<std::size_t>(nterms[g]) * F.nent * F.nq;
            }
            mo += nmark[g];
            to += nterms[g];
        }
    });
}

/// The program at points with their normals (ds terms).
int orc_integrand_eval_n(int ninstr, const int *code, int nreg, const double *consts, long npts, const double *xs,
                         const double *ys, const double *nxs, const double *nys, double *out)
{
    return guard([&] {
        std::vector<double> reg;
        for (long i = 0; i < npts; ++i) out[i] = run_program(ninstr, code, nreg, consts, xs[i], ys[i], reg, nxs[i], nys[i]);
    });
}

int orc_spmv(int n, const int *ip, const int *ix, const double *v, const double *x, double *y)
{
    return guard([&] { spmv(n, ip, ix, v, x, y); });
}

double orc_ddot(long n, const double *x, const double *y) { return ddot(x, y, static_cast<std::size_t>(n)); }

int orc_rcm(int n, const int *ip, const int *ix, int *perm)
{
    return guard([&] { std::vector<int> p = rcm(n, ip, ix); std::copy(p.begin(), p.end(), perm); });
}

/// The ILU(0) factors on P A P^T (perm may be the identity when ordering = 0): the permuted
/// pattern, the factor values (L strictly below the diagonal, unit diagonal implied; U on and
/// above it), and the smallest pivot.  ip_out has n+1 entries, ix_out and lu_out nnz.
int orc_ilu0_factor(int n, const int *ip, const int *ix, const double *v, int ordering, int *perm,
                    int *ip_out, int *ix_out, double *lu_out, double *min_pivot)
{
    return guard([&] {
        Precond M(make_csr(n, ip, ix, v), 2, ordering, 1.0);
        for (int k = 0; k < n; ++k) perm[k] = M.perm.empty() ? k : M.perm[k];
        std::copy(M.LU.ip.begin(), M.LU.ip.end(), ip_out);
        std::copy(M.LU.ix.begin(), M.LU.ix.end(), ix_out);
        std::copy(M.LU.v.begin(), M.LU.v.end(), lu_out);
        *min_pivot = M.min_pivot;
    });
}

int orc_precond_apply(int n, const int *ip, const int *ix, const double *v, int kind, int ordering, double omega,
                      const double *r, double *z, double *min_pivot)
{
    return guard([&] {
        Precond M(make_csr(n, ip, ix, v), kind, ordering, omega);
        M.apply(r, z);
        if (min_pivot) *min_pivot = M.min_pivot;
    });
}

/// CG as python/src/femd_bindings.cpp:799-833 dispatches it: kinds 0 (none) and 1 (Jacobi)
/// take cg_jacobi, 2 (ILU(0)) and 3 (SSOR) the general loop.  info: converged, iterations, breakdown.
int orc_cg(int n, const int *ip, const int *ix, const double *v, int pc, int ordering, double omega,
           const double *b, double *x, int max_iter, double tol, int *info, double *res)
{
    return guard([&] {
        CSR A = make_csr(n, ip, ix, v);
        std::vector<double> xv(x, x + n), bv(b, b + n);
        CGOut o;
        if (pc == 0) o = cg_jacobi(A, nullptr, xv, bv, max_iter, tol);
        else
        {
            Precond M(A, pc, ordering, omega);
            if (pc == 1) o = cg_jacobi(A, M.inv.data(), xv, bv, max_iter, tol);
            else o = cg_general(A, M, xv, bv, max_iter, tol);
        }
        std::copy(xv.begin(), xv.end(), x);
        info[0] = o.converged; info[1] = o.iterations; info[2] = o.breakdown;
        *res = o.residual;
    });
}

/// krylov/adapters.hpp:116-147.  res: true residual, preconditioned residual.  info: converged, iterations.
int orc_krylov(int n, const int *ip, const int *ix, const double *v, int pc, int ordering, double omega,
               const double *b, double *x, int use_lgmres, int restart, int max_iter, double tol, int k_aug, int *info, double *res)
{
    return guard([&] {
        CSR A = make_csr(n, ip, ix, v);
        std::unique_ptr<Precond> M;
        if (pc != 0) M = std::make_unique<Precond>(A, pc, ordering, omega);
        auto Aop = [&A](const std::vector<double> &in, std::vector<double> &out) {
            out.resize(static_cast<std::size_t>(A.n));
            spmv(A.n, A.ip.data(), A.ix.data(), A.v.data(), in.data(), out.data());
        };
        auto Mop = [&M](const std::vector<double> &in, std::vector<double> &out) {
            out.resize(in.size());
            if (M) M->apply(in.data(), out.data()); else out = in;
        };
        std::vector<double> xv(x, x + n), bv(b, b + n);
        const double nb = snorm(bv);
        if (nb == 0.0)
        {
            std::fill(x, x + n, 0.0);
            info[0] = 1; info[1] = 0; res[0] = res[1] = 0.0;
            return;
        }
        std::vector<double> Mb;
        Mop(bv, Mb);
        const double nMb = snorm(Mb), scaled = nMb > 0.0 ? tol * nMb / nb : tol;
        GOut g = use_lgmres ? lgmres(Aop, Mop, xv, bv, restart, k_aug, max_iter, scaled) : gmres(Aop, Mop, xv, bv, restart, max_iter, scaled);
        std::vector<double> Ax, r(static_cast<std::size_t>(n)), Mr;
        Aop(xv, Ax);
        for (int i = 0; i < n; ++i) r[i] = bv[i] - Ax[i];
        Mop(r, Mr);
        std::copy(xv.begin(), xv.end(), x);
        res[0] = snorm(r) / nb;
        res[1] = nMb > 0.0 ? snorm(Mr) / nMb : 0.0;
        info[0] = res[1] <= tol * (1.0 + 1e-8) || (g.converged && res[1] <= 10.0 * tol);
        info[1] = g.iterations;
    });
}

} // extern "C"
