// host.cpp -- the host-side setup of the MI355X path (see host.hpp).
//
// Where the reference's bits depend on an operation order (node coordinates, the basis
// tables, quadrature points and weights, per-cell Jacobians), the same floating-point
// operations are performed in the same order; the integer structure (edges, numbering,
// pattern, gather lists) is built with sorts over flat arrays instead of hash maps and
// per-row containers.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <numeric>
#include <utility>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace femd_amd {

namespace {
inline int64_t edge_key(int a, int b)
{
    if (a > b) std::swap(a, b);
    return (static_cast<int64_t>(a) << 32) | static_cast<uint32_t>(b);
}
[[noreturn]] void bad(const std::string &m) { throw Error(1, m); }
} // namespace

int MeshData::marker_of(int a, int b) const
{
    const int64_t k = edge_key(a, b);
    auto it = std::lower_bound(seg_key.begin(), seg_key.end(), k);
    if (it == seg_key.end() || *it != k) return -1;
    return smark[seg_of_key[it - seg_key.begin()]];
}

// ------------------------------------------------------------------------------------ meshes
// mesh/structured.hpp:99-221.  Half-edges are matched by sorting (key, slot) pairs: the two
// sides of an interior edge are adjacent after the sort.
MeshData mesh_from_arrays(std::vector<double> xy, std::vector<int> tri, const std::vector<int> &segs,
                          const std::vector<int> &smarks, int marker)
{
    const int np = static_cast<int>(xy.size() / 2), nt = static_cast<int>(tri.size() / 3);
    if (np < 3 || nt < 1) bad("mesh_from_arrays: need at least 3 points and 1 triangle");
    if (!smarks.empty() && smarks.size() * 2 != segs.size()) bad("mesh_from_arrays: one marker per segment");
    std::vector<char> used(static_cast<std::size_t>(np), 0);
    for (int t = 0; t < nt; ++t)
    {
        int *T = &tri[3 * static_cast<std::size_t>(t)];
        for (int k = 0; k < 3; ++k)
        {
            if (T[k] < 0 || T[k] >= np) bad("mesh_from_arrays: triangle " + std::to_string(t) + " has a vertex index out of range");
            used[T[k]] = 1;
        }
        if (T[0] == T[1] || T[1] == T[2] || T[0] == T[2]) bad("mesh_from_arrays: triangle " + std::to_string(t) + " repeats a vertex");
        const double *a = &xy[2 * T[0]], *b = &xy[2 * T[1]], *c = &xy[2 * T[2]];
        const double o = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
        if (o == 0.0) bad("mesh_from_arrays: triangle " + std::to_string(t) + " is degenerate");
        if (o < 0.0) std::swap(T[1], T[2]);
    }
    for (int v = 0; v < np; ++v)
        if (!used[v]) bad("mesh_from_arrays: point " + std::to_string(v) + " is not a vertex of any triangle");

    MeshData M;
    M.xy = std::move(xy);
    M.tri = std::move(tri);
    M.nbr.assign(3 * static_cast<std::size_t>(nt), -1);
    std::vector<std::pair<int64_t, int>> half(3 * static_cast<std::size_t>(nt));
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < 3; ++i)
            half[3 * t + i] = {edge_key(M.tri[3 * t + (i + 1) % 3], M.tri[3 * t + (i + 2) % 3]), 3 * t + i};
    std::sort(half.begin(), half.end());
    for (std::size_t h = 0; h < half.size();)
    {
        std::size_t e = h + 1;
        while (e < half.size() && half[e].first == half[h].first) ++e;
        if (e - h > 2) bad("mesh_from_arrays: an edge is shared by more than two triangles");
        if (e - h == 2)
        {
            const int s0 = half[h].second, s1 = half[h + 1].second;
            const int a0 = M.tri[s0 - s0 % 3 + (s0 % 3 + 1) % 3], a1 = M.tri[s1 - s1 % 3 + (s1 % 3 + 1) % 3];
            if (a0 == a1) bad("mesh_from_arrays: two triangles lie on the same side of an edge (the mesh overlaps or folds)");
            M.nbr[s0] = s1 / 3;
            M.nbr[s1] = s0 / 3;
        }
        h = e;
    }
    // boundary edges, in triangle order, directed as in their triangle
    std::vector<std::array<int, 2>> bedge;
    std::vector<int> out_start(static_cast<std::size_t>(np) + 1, 0);
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < 3; ++i)
            if (M.nbr[3 * t + i] < 0) bedge.push_back({M.tri[3 * t + (i + 1) % 3], M.tri[3 * t + (i + 2) % 3]});
    for (auto &e : bedge) ++out_start[e[0] + 1];
    for (int v = 0; v < np; ++v) out_start[v + 1] += out_start[v];
    std::vector<int> out_list(bedge.size()), fill(out_start.begin(), out_start.end() - 1);
    for (std::size_t e = 0; e < bedge.size(); ++e) out_list[fill[bedge[e][0]]++] = static_cast<int>(e);

    // segments: the user's first (their markers win), then unlisted boundary edges loop by loop
    std::vector<std::pair<int64_t, int>> keys;           // (edge key, segment index), kept sorted
    auto find_seg = [&](int64_t k) {
        auto it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(k, -1));
        return (it != keys.end() && it->first == k) ? it->second : -1;
    };
    auto add_seg = [&](int a, int b, int m) {
        const int64_t k = edge_key(a, b);
        const int s = find_seg(k);
        if (s >= 0) { M.smark[s] = m; return; }
        const int idx = static_cast<int>(M.smark.size());
        M.seg.push_back(a); M.seg.push_back(b); M.smark.push_back(m);
        keys.insert(std::upper_bound(keys.begin(), keys.end(), std::make_pair(k, idx)), std::make_pair(k, idx));
    };
    const int ns = static_cast<int>(segs.size() / 2);
    for (int s = 0; s < ns; ++s)
    {
        const int a = segs[2 * s], b = segs[2 * s + 1];
        const int64_t k = edge_key(a, b);
        if (a < 0 || a >= np || b < 0 || b >= np ||
            !std::binary_search(half.begin(), half.end(), std::make_pair(k, -1), [](const auto &x, const auto &y) { return x.first < y.first; }))
            bad("mesh_from_arrays: segment " + std::to_string(s) + " is not an edge of the mesh");
        add_seg(a, b, smarks.empty() ? marker : smarks[s]);
    }
    std::vector<char> done(bedge.size(), 0);
    for (int start = 0; start < np; ++start)
        for (int q = out_start[start]; q < out_start[start + 1]; ++q)
        {
            int e = out_list[q];
            while (e >= 0 && !done[e])
            {
                done[e] = 1;
                const int a = bedge[e][0], b = bedge[e][1];
                if (find_seg(edge_key(a, b)) < 0) add_seg(a, b, marker);
                int next = -1;
                for (int r = out_start[b]; r < out_start[b + 1]; ++r)
                    if (!done[out_list[r]]) { next = out_list[r]; break; }
                e = next;
            }
        }
    M.seg_key.resize(keys.size()); M.seg_of_key.resize(keys.size());
    for (std::size_t i = 0; i < keys.size(); ++i) { M.seg_key[i] = keys[i].first; M.seg_of_key[i] = keys[i].second; }
    return M;
}

// mesh/structured.hpp:260-351 (grid coordinates, the four diagonal patterns, sides as segments)
MeshData mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, int diag, const int markers[4])
{
    if (!(x1 > x0 && y1 > y0)) bad("rectangle_mesh: need x1 > x0 and y1 > y0");
    if (nx < 1 || ny < 1) bad("rectangle_mesh: need nx, ny >= 1");
    if (diag < 0 || diag > 3) bad("rectangle_mesh: unknown diagonal");
    const int W = nx + 1;
    std::vector<double> g(2 * static_cast<std::size_t>(W) * (ny + 1));
    for (int j = 0; j <= ny; ++j)
    {
        const double y = j == ny ? y1 : y0 + (y1 - y0) * j / ny;
        for (int i = 0; i <= nx; ++i)
        {
            g[2 * (static_cast<std::size_t>(j) * W + i)] = i == nx ? x1 : x0 + (x1 - x0) * i / nx;
            g[2 * (static_cast<std::size_t>(j) * W + i) + 1] = y;
        }
    }
    const int nc0 = W * (ny + 1);
    std::vector<int> tri;
    tri.reserve(static_cast<std::size_t>(nx) * ny * (diag == 3 ? 12 : 6));
    if (diag == 3)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
            {
                const int a = j * W + i, b = a + 1, c = a + W + 1, d = a + W;
                g.push_back(0.25 * (g[2 * a] + g[2 * b] + g[2 * c] + g[2 * d]));
                g.push_back(0.25 * (g[2 * a + 1] + g[2 * b + 1] + g[2 * c + 1] + g[2 * d + 1]));
            }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
            const int a = j * W + i, b = a + 1, c = a + W + 1, d = a + W;
            if (diag == 3)
            {
                const int m = nc0 + j * nx + i;
                tri.insert(tri.end(), {a, b, m, b, c, m, c, d, m, d, a, m});
            }
            else if (diag == 0 || (diag == 2 && (i + j) % 2 == 0)) tri.insert(tri.end(), {a, b, c, a, c, d});
            else tri.insert(tri.end(), {a, b, d, b, c, d});
        }
    std::vector<int> segs, smk;
    auto side = [&](int a, int b, int m) { segs.push_back(a); segs.push_back(b); smk.push_back(m); };
    for (int i = 0; i < nx; ++i) side(i, i + 1, markers[0]);
    for (int j = 0; j < ny; ++j) side(j * W + nx, (j + 1) * W + nx, markers[1]);
    for (int i = nx; i > 0; --i) side(ny * W + i, ny * W + i - 1, markers[2]);
    for (int j = ny; j > 0; --j) side(j * W, (j - 1) * W, markers[3]);
    return mesh_from_arrays(std::move(g), std::move(tri), segs, smk, markers[0]);
}

// mesh/quad_mesh.hpp:50-103.  Clockwise quads are re-oriented by swapping vertices 1 and 3; every
// quad must be strictly convex; neighbours by sorting (edge key, slot) pairs; the listed
// segments, then every other boundary edge in quad order with `marker`.
MeshData quad_mesh_from_arrays(std::vector<double> xy, std::vector<int> quads, const std::vector<int> &segs,
                               const std::vector<int> &smarks, int marker)
{
    const int np = static_cast<int>(xy.size() / 2), nq = static_cast<int>(quads.size() / 4);
    if (np < 4 || nq < 1) bad("quad_mesh_from_arrays: need at least 4 points and 1 quad");
    if (!smarks.empty() && smarks.size() * 2 != segs.size()) bad("quad_mesh_from_arrays: one marker per segment");
    auto P = [&](int v, int d) { return xy[2 * static_cast<std::size_t>(v) + d]; };
    for (int t = 0; t < nq; ++t)
    {
        int *Q = &quads[4 * static_cast<std::size_t>(t)];
        for (int k = 0; k < 4; ++k)
            if (Q[k] < 0 || Q[k] >= np) bad("quad_mesh_from_arrays: quad " + std::to_string(t) + " has a vertex out of range");
        double s = 0.0;
        for (int i = 0; i < 4; ++i) s += P(Q[i], 0) * P(Q[(i + 1) % 4], 1) - P(Q[(i + 1) % 4], 0) * P(Q[i], 1);
        if (0.5 * s < 0.0) std::swap(Q[1], Q[3]);
        for (int i = 0; i < 4; ++i)
        {
            const int a = Q[(i + 3) % 4], b = Q[i], c = Q[(i + 1) % 4];
            const double o = (P(b, 0) - P(a, 0)) * (P(c, 1) - P(a, 1)) - (P(b, 1) - P(a, 1)) * (P(c, 0) - P(a, 0));
            if (!(o > 0.0)) bad("quad_mesh_from_arrays: quad " + std::to_string(t) + " is degenerate or not convex");
        }
    }
    MeshData M;
    M.nv = 4;
    M.xy = std::move(xy);
    M.tri = std::move(quads);
    M.nbr.assign(4 * static_cast<std::size_t>(nq), -1);
    std::vector<std::pair<int64_t, int>> half(4 * static_cast<std::size_t>(nq));
    for (int t = 0; t < nq; ++t)
        for (int i = 0; i < 4; ++i)
        {
            int a, b;
            M.edge(t, i, a, b);
            half[4 * static_cast<std::size_t>(t) + i] = {edge_key(a, b), 4 * t + i};
        }
    std::sort(half.begin(), half.end());
    for (std::size_t h = 0; h < half.size();)
    {
        std::size_t e = h;
        while (e < half.size() && half[e].first == half[h].first) ++e;
        if (e - h > 2) bad("quad_mesh_from_arrays: an edge is shared by more than two quads");
        if (e - h == 2)
        {
            const int s0 = half[h].second, s1 = half[h + 1].second;
            M.nbr[s0] = s1 / 4;
            M.nbr[s1] = s0 / 4;
        }
        h = e;
    }
    std::vector<std::pair<int64_t, int>> keys;
    auto add = [&](int a, int b, int m) {
        M.seg.push_back(a); M.seg.push_back(b); M.smark.push_back(m);
        keys.push_back({edge_key(a, b), static_cast<int>(M.smark.size()) - 1});
    };
    std::vector<int64_t> listed;
    for (std::size_t s2 = 0; 2 * s2 < segs.size(); ++s2)
    {
        const int64_t k = edge_key(segs[2 * s2], segs[2 * s2 + 1]);
        if (!std::binary_search(half.begin(), half.end(), std::make_pair(k, -1),
                                [](const std::pair<int64_t, int> &x, const std::pair<int64_t, int> &y) { return x.first < y.first; }))
            bad("quad_mesh_from_arrays: segment " + std::to_string(s2) + " is not an edge of the mesh");
        add(segs[2 * s2], segs[2 * s2 + 1], smarks.empty() ? marker : smarks[s2]);
        listed.push_back(k);
    }
    std::sort(listed.begin(), listed.end());
    for (int t = 0; t < nq; ++t)
        for (int i = 0; i < 4; ++i)
            if (M.nbr[4 * static_cast<std::size_t>(t) + i] < 0)
            {
                int a, b;
                M.edge(t, i, a, b);
                if (!std::binary_search(listed.begin(), listed.end(), edge_key(a, b))) add(a, b, marker);
            }
    std::sort(keys.begin(), keys.end());
    M.seg_key.resize(keys.size()); M.seg_of_key.resize(keys.size());
    for (std::size_t i = 0; i < keys.size(); ++i) { M.seg_key[i] = keys[i].first; M.seg_of_key[i] = keys[i].second; }
    return M;
}

MeshData quad_mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, const int markers[4])
{
    const MeshData T = mesh_rectangle(x0, y0, x1, y1, nx, ny, 0, markers);
    const int W = nx + 1;
    std::vector<double> g(T.xy.begin(), T.xy.begin() + 2 * static_cast<std::size_t>(W) * (ny + 1));
    std::vector<int> q;
    q.reserve(4 * static_cast<std::size_t>(nx) * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
            const int v = j * W + i;
            q.insert(q.end(), {v, v + 1, v + W + 1, v + W});
        }
    return quad_mesh_from_arrays(std::move(g), std::move(q), T.seg, T.smark, markers[0]);
}

// -------------------------------------------------------------------------- reference element
namespace {
// Legendre P_n, P_n', P_n'' at x for n = 0..p by the three-term recurrence with derivatives
// n P_n^(m) = (2n-1)(x P_{n-1}^(m) + m P_{n-1}^(m-1)) - (n-1) P_{n-2}^(m) (fe/dg_space.hpp:38-58),
// in that evaluation order so the Newton iterates below carry the reference's bits.
void legendre_d2(int p, double x, double *rows)
{
    const int nl = p + 1;
    for (int m = 0; m < 3; ++m)
    {
        double *R = rows + m * nl;
        const double *D = m > 0 ? rows + (m - 1) * nl : nullptr;
        R[0] = m == 0 ? 1.0 : 0.0;
        for (int n = 1; n <= p; ++n)
        {
            const double lower = D ? D[n - 1] : 0.0, prev2 = n >= 2 ? R[n - 2] : 0.0;
            R[n] = ((2 * n - 1) * (x * R[n - 1] + m * lower) - (n - 1) * prev2) / n;
        }
    }
}

// The p+1 Gauss-Lobatto points on [-1, 1]: the ends and the roots of P_p', each by Newton
// (step P_p'/P_p'') from the Chebyshev-Lobatto guess -cos(pi i / p), at most 100 steps, stopping
// after a step below 1e-16 (fe/dg_space.hpp:60-80).
std::vector<double> lobatto(int p)
{
    std::vector<double> x(static_cast<std::size_t>(p) + 1), rows(3 * (static_cast<std::size_t>(p) + 1));
    x[0] = -1.0;
    x[p] = 1.0;
    for (int i = 1; i < p; ++i)
    {
        double z = -std::cos(3.14159265358979323846 * i / p);
        for (int it = 0; it < 100; ++it)
        {
            legendre_d2(p, z, rows.data());
            const double step = rows[(p + 1) + p] / rows[2 * (p + 1) + p];
            z -= step;
            if (std::fabs(step) < 1e-16) break;
        }
        x[i] = z;
    }
    return x;
}

// Hesthaven's 1D warp at r: the Lobatto-minus-uniform node shift interpolated on the uniform
// nodes, divided by 1 - r^2, zero within 1e-10 of the ends (simplex_basis.hpp:150-166).
double warp(int N, double r, const std::vector<double> &lgl)
{
    double w = 0.0;
    for (int i = 0; i <= N; ++i)
    {
        const double ri = -1.0 + 2.0 * i / N;
        double ell = 1.0;
        for (int j = 0; j <= N; ++j)
            if (j != i) ell *= (r - (-1.0 + 2.0 * j / N)) / (ri - (-1.0 + 2.0 * j / N));
        w += ell * (lgl[i] - ri);
    }
    if (std::abs(r) < 1.0 - 1e-10) return w / (1.0 - r * r);
    return 0.0;
}

// Warburton's warp-and-blend nodes of degree N (simplex_basis.hpp:168-193): the lattice on the
// equilateral triangle, each edge's warp blended inwards with the per-degree optimum alpha (5/3
// from degree 16), then mapped to the reference triangle (xi, eta).
void warp_blend(int N, std::vector<double> &xs, std::vector<double> &ys)
{
    static const double alpha_opt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                         1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
    const double al = N < 16 ? alpha_opt[N - 1] : 5.0 / 3.0;
    const std::vector<double> lgl = lobatto(N);
    const double pi = 3.14159265358979323846, s3 = std::sqrt(3.0);
    const double c2 = std::cos(2.0 * pi / 3.0), c4 = std::cos(4.0 * pi / 3.0);
    const double s2 = std::sin(2.0 * pi / 3.0), s4 = std::sin(4.0 * pi / 3.0);
    for (int n = 0; n <= N; ++n)
        for (int m = 0; m <= N - n; ++m)
        {
            const double L1 = static_cast<double>(n) / N, L3 = static_cast<double>(m) / N, L2 = 1.0 - L1 - L3;
            double x = -L2 + L3, y = (-L2 - L3 + 2.0 * L1) / s3;
            const double w1 = 4.0 * L2 * L3 * warp(N, L3 - L2, lgl) * (1.0 + (al * L1) * (al * L1));
            const double w2 = 4.0 * L1 * L3 * warp(N, L1 - L3, lgl) * (1.0 + (al * L2) * (al * L2));
            const double w3 = 4.0 * L1 * L2 * warp(N, L2 - L1, lgl) * (1.0 + (al * L3) * (al * L3));
            x += w1 + c2 * w2 + c4 * w3;
            y += s2 * w2 + s4 * w3;
            const double l1 = (s3 * y + 1.0) / 3.0, l2 = (-3.0 * x - s3 * y + 2.0) / 6.0, l3 = (3.0 * x - s3 * y + 2.0) / 6.0;
            xs.push_back(0.5 * ((-l2 + l3 - l1) + 1.0));
            ys.push_back(0.5 * ((-l2 - l3 + l1) + 1.0));
        }
}

// Orthonormal Jacobi P_0..P_N^{(a,b)} (simplex_basis.hpp:38-55, Hesthaven's JacobiP).
void jacobi_table(double x, double al, double be, int N, double *P)
{
    const double ab = al + be;
    const double g0 = std::pow(2.0, ab + 1.0) / (ab + 1.0) * std::tgamma(al + 1.0) * std::tgamma(be + 1.0) / std::tgamma(ab + 1.0);
    P[0] = 1.0 / std::sqrt(g0);
    if (N == 0) return;
    const double g1 = (al + 1.0) * (be + 1.0) / (ab + 3.0) * g0;
    P[1] = ((ab + 2.0) * x / 2.0 + (al - be) / 2.0) / std::sqrt(g1);
    double aprev = 2.0 / (2.0 + ab) * std::sqrt((al + 1.0) * (be + 1.0) / (ab + 3.0));
    for (int i = 1; i < N; ++i)
    {
        const double h = 2.0 * i + ab;
        const double an = 2.0 / (h + 2.0) * std::sqrt((i + 1.0) * (i + 1.0 + ab) * (i + 1.0 + al) * (i + 1.0 + be) / (h + 1.0) / (h + 3.0));
        const double bn = -(al * al - be * be) / h / (h + 2.0);
        P[i + 1] = 1.0 / an * (-aprev * P[i - 1] + (x - bn) * P[i]);
        aprev = an;
    }
}

double jacobi_deriv(double x, double al, double be, int n)
{
    if (n == 0) return 0.0;
    double P[32];
    jacobi_table(x, al + 1.0, be + 1.0, n - 1, P);
    return std::sqrt(n * (n + al + be + 1.0)) * P[n - 1];
}

// Dubiner modes and their xi/eta derivatives (simplex_basis.hpp:77-114).
void modes(int k, double xi, double eta, double *v, double *dx, double *dy)
{
    const double r = 2.0 * xi - 1.0, s = 2.0 * eta - 1.0;
    const double a = (std::abs(1.0 - s) > 1e-14) ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0, b = s;
    double Pa[32], dPa[32], Pb[32];
    jacobi_table(a, 0.0, 0.0, k, Pa);
    for (int i = 0; i <= k; ++i) dPa[i] = jacobi_deriv(a, 0.0, 0.0, i);
    const double hb = 0.5 * (1.0 - b);
    int m = 0;
    for (int i = 0; i <= k; ++i)
    {
        jacobi_table(b, 2.0 * i + 1.0, 0.0, k - i, Pb);
        const double hbi = std::pow(hb, i), hbm = i > 0 ? std::pow(hb, i - 1) : 0.0, c = std::pow(2.0, i + 0.5);
        for (int j = 0; j <= k - i; ++j, ++m)
        {
            v[m] = c * Pa[i] * Pb[j] * hbi;
            if (!dx) continue;
            double dr = dPa[i] * Pb[j];
            if (i > 0) dr *= hbm;
            double ds = dPa[i] * Pb[j] * 0.5 * (1.0 + a);
            if (i > 0) ds *= hbm;
            double t = jacobi_deriv(b, 2.0 * i + 1.0, 0.0, j) * hbi;
            if (i > 0) t -= 0.5 * i * Pb[j] * hbm;
            ds += Pa[i] * t;
            dr *= c; ds *= c;
            dx[m] = 2.0 * dr;
            dy[m] = 2.0 * ds;
        }
    }
}
} // namespace

RefElement::RefElement(int degree, bool equispaced) : k(degree), n((degree + 1) * (degree + 2) / 2)
{
    if (k < 1 || k > 20) bad("LagrangeSpace: degree must be in 1..20");
    static const double vx[3] = {0.0, 1.0, 0.0}, vy[3] = {0.0, 0.0, 1.0};
    nxi.assign(vx, vx + 3); neta.assign(vy, vy + 3);
    std::vector<double> z(static_cast<std::size_t>(k) + 1);
    if (equispaced)
        for (int j = 0; j <= k; ++j) z[j] = -1.0 + 2.0 * j / k;
    else
        z = lobatto(k);
    for (int j = 0; j <= k; ++j) z[j] = 0.5 * (z[j] - z[k - j]);          // symmetrized in place, as the reference does
    for (int e = 0; e < 3; ++e)
        for (int j = 1; j < k; ++j)
        {
            const double t = 0.5 * (1.0 + z[j]);
            nxi.push_back((1.0 - t) * vx[(e + 1) % 3] + t * vx[(e + 2) % 3]);
            neta.push_back((1.0 - t) * vy[(e + 1) % 3] + t * vy[(e + 2) % 3]);
        }
    if (equispaced)
    {
        for (int j = 1; j < k; ++j)
            for (int i = 1; i + j < k; ++i) { nxi.push_back(static_cast<double>(i) / k); neta.push_back(static_cast<double>(j) / k); }
    }
    else if (k >= 3)
    {
        // the warp-and-blend points strictly inside (simplex_basis.hpp:274-282)
        std::vector<double> px, py;
        warp_blend(k, px, py);
        for (std::size_t p = 0; p < px.size(); ++p)
            if (px[p] > 1e-10 && py[p] > 1e-10 && 1.0 - px[p] - py[p] > 1e-10) { nxi.push_back(px[p]); neta.push_back(py[p]); }
    }
    if (static_cast<int>(nxi.size()) != n) bad("reference element: wrong node count");
    // Vandermonde V_im = psi_m(node_i), inverted by Gauss-Jordan with partial pivoting
    std::vector<double> A(static_cast<std::size_t>(n) * n), X(static_cast<std::size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) modes(k, nxi[i], neta[i], &A[static_cast<std::size_t>(i) * n], nullptr, nullptr);
    for (int i = 0; i < n; ++i) X[static_cast<std::size_t>(i) * n + i] = 1.0;
    for (int c = 0; c < n; ++c)
    {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(A[r * n + c]) > std::abs(A[piv * n + c])) piv = r;
        if (A[piv * n + c] == 0.0) bad("reference element: singular Vandermonde matrix");
        if (piv != c)
            for (int j = 0; j < n; ++j) { std::swap(A[piv * n + j], A[c * n + j]); std::swap(X[piv * n + j], X[c * n + j]); }
        const double d = A[c * n + c];
        for (int j = 0; j < n; ++j) { A[c * n + j] /= d; X[c * n + j] /= d; }
        for (int r = 0; r < n; ++r)
        {
            const double f = A[r * n + c];
            if (r == c || f == 0.0) continue;
            for (int j = 0; j < n; ++j) { A[r * n + j] -= f * A[c * n + j]; X[r * n + j] -= f * X[c * n + j]; }
        }
    }
    vinv = std::move(X);
}

void RefElement::tabulate(double xi, double eta, double *out) const
{
    std::vector<double> psi(3 * static_cast<std::size_t>(n));
    modes(k, xi, eta, psi.data(), psi.data() + n, psi.data() + 2 * n);
    for (int r = 0; r < 3; ++r)
    {
        double *o = out + r * n;
        std::fill(o, o + n, 0.0);
        for (int m = 0; m < n; ++m)
        {
            const double a = psi[r * n + m];
            if (a == 0.0) continue;
            for (int l = 0; l < n; ++l) o[l] += a * vinv[m * n + l];
        }
    }
}

// ----------------------------------------------------------------------------- quadrature
namespace {
// quadrature/gauss_legendre.hpp:30-137
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0); w.assign(n, 0.0);
    switch (n)
    {
    case 1: x = {0.0}; w = {2.0}; return;
    case 2: x = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)}; w = {1.0, 1.0}; return;
    case 3: x = {-std::sqrt(3.0 / 5.0), 0.0, std::sqrt(3.0 / 5.0)}; w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0}; return;
    case 4:
    {
        const double a = std::sqrt((3.0 + 2.0 * std::sqrt(6.0 / 5.0)) / 7.0), b = std::sqrt((3.0 - 2.0 * std::sqrt(6.0 / 5.0)) / 7.0);
        const double wa = (18.0 - std::sqrt(30.0)) / 36.0, wb = (18.0 + std::sqrt(30.0)) / 36.0;
        x = {-a, -b, b, a}; w = {wa, wb, wb, wa};
        return;
    }
    case 5:
    {
        const double a = std::sqrt(5.0 + 2.0 * std::sqrt(10.0 / 7.0)) / 3.0, b = std::sqrt(5.0 - 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        const double wa = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0, wb = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        x = {-a, -b, 0.0, b, a}; w = {wa, wb, 128.0 / 225.0, wb, wa};
        return;
    }
    default: break;
    }
    const double PI = 3.14159265358979323846;
    auto legendre = [n](double z, double &p0, double &p1) {
        p0 = 1.0; p1 = z;
        for (int k = 2; k <= n; ++k) { const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k; p0 = p1; p1 = p2; }
    };
    for (int i = 0; i < n; ++i)
    {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5)), p0, p1;
        for (int it = 0; it < 100; ++it)
        {
            legendre(z, p0, p1);
            const double dz = -p1 / (n * (z * p1 - p0) / (z * z - 1.0));
            z += dz;
            if (std::abs(dz) <= 1e-15) break;
        }
        legendre(z, p0, p1);
        const double dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}
} // namespace

// quadrature/triangle.hpp:22-57: 1, 3 and Radon's 7 points, then the collapsed Gauss rule
QuadRule triangle_rule(int d)
{
    QuadRule Q;
    auto add = [&Q](double a, double b, double c) { Q.xi.push_back(a); Q.eta.push_back(b); Q.w.push_back(c); };
    if (d <= 1) add(1.0 / 3.0, 1.0 / 3.0, 0.5);
    else if (d == 2) { add(1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0); add(2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0); add(1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0); }
    else if (d <= 5)
    {
        const double s15 = std::sqrt(15.0);
        const double a1 = (6.0 - s15) / 21.0, w1 = (155.0 - s15) / 2400.0, a2 = (6.0 + s15) / 21.0, w2 = (155.0 + s15) / 2400.0;
        add(1.0 / 3.0, 1.0 / 3.0, 9.0 / 80.0);
        add(a1, a1, w1); add(1.0 - 2.0 * a1, a1, w1); add(a1, 1.0 - 2.0 * a1, w1);
        add(a2, a2, w2); add(1.0 - 2.0 * a2, a2, w2); add(a2, 1.0 - 2.0 * a2, w2);
    }
    else
    {
        std::vector<double> ux, uw, vx, vw;
        gauss_legendre((d + 2) / 2, ux, uw);
        gauss_legendre((d + 3) / 2, vx, vw);
        for (std::size_t j = 0; j < vx.size(); ++j)
        {
            const double v = 0.5 * (vx[j] + 1.0), wv = 0.5 * vw[j];
            for (std::size_t i = 0; i < ux.size(); ++i)
            {
                const double u = 0.5 * (ux[i] + 1.0), wu = 0.5 * uw[i];
                add(u * (1.0 - v), v, wu * wv * (1.0 - v));
            }
        }
    }
    return Q;
}

RefQuad::RefQuad(int degree, bool equispaced) : k(degree), n((degree + 1) * (degree + 1))
{
    if (k < 1 || k > 20) bad("LagrangeSpaceQ: degree must be in 1..20");
    std::vector<double> z;
    if (equispaced)
        for (int j = 0; j <= k; ++j) z.push_back(-1.0 + 2.0 * j / k);
    else
        z = lobatto(k);
    t.resize(static_cast<std::size_t>(k) + 1);
    for (int j = 0; j <= k; ++j) t[j] = 0.5 * (1.0 + 0.5 * (z[j] - z[k - j]));
    t[0] = 0.0;
    t[k] = 1.0;
    w.assign(static_cast<std::size_t>(k) + 1, 1.0);
    for (int j = 0; j <= k; ++j)
        for (int m = 0; m <= k; ++m)
            if (m != j) w[j] /= (t[j] - t[m]);
    auto add = [&](int i, int j) { ii.push_back(i); jj.push_back(j); };
    add(0, 0); add(k, 0); add(k, k); add(0, k);
    for (int j = 1; j < k; ++j) add(j, 0);           // edge 0, bottom
    for (int j = 1; j < k; ++j) add(k, j);           // edge 1, right
    for (int j = 1; j < k; ++j) add(k - j, k);       // edge 2, top
    for (int j = 1; j < k; ++j) add(0, k - j);       // edge 3, left
    for (int j = 1; j < k; ++j)
        for (int i = 1; i < k; ++i) add(i, j);
    for (int l = 0; l < n; ++l) { nxi.push_back(t[ii[l]]); neta.push_back(t[jj[l]]); }
}

void RefQuad::tabulate(double xi, double eta, double *out) const
{
    std::vector<double> a(4 * (static_cast<std::size_t>(k) + 1));
    double *L[2] = {a.data(), a.data() + 2 * (k + 1)}, *dL[2] = {a.data() + (k + 1), a.data() + 3 * (k + 1)};
    const double x[2] = {xi, eta};
    for (int d = 0; d < 2; ++d)
        for (int j = 0; j <= k; ++j)
        {
            double p = 1.0, dp = 0.0;
            for (int m = 0; m <= k; ++m)
            {
                if (m == j) continue;
                dp = dp * (x[d] - t[m]) + p;       // product rule, one factor differentiated at a time
                p *= (x[d] - t[m]);
            }
            L[d][j] = w[j] * p;
            dL[d][j] = w[j] * dp;
        }
    for (int l = 0; l < n; ++l)
    {
        out[l] = L[0][ii[l]] * L[1][jj[l]];
        out[n + l] = dL[0][ii[l]] * L[1][jj[l]];
        out[2 * n + l] = L[0][ii[l]] * dL[1][jj[l]];
    }
}

QuadRule square_rule(int d)
{
    std::vector<double> g, gw;
    gauss_legendre(std::max(1, (d + 3) / 2), g, gw);
    QuadRule Q;
    const int m = static_cast<int>(g.size());
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i)
        {
            Q.xi.push_back(0.5 * (g[i] + 1.0));
            Q.eta.push_back(0.5 * (g[j] + 1.0));
            Q.w.push_back(0.25 * gw[i] * gw[j]);
        }
    return Q;
}

// ---------------------------------------------------------------------------------- spaces
SpaceData build_space(const MeshData &M, int k, std::vector<int> dirichlet, bool all_exterior, bool equispaced)
{
    std::sort(dirichlet.begin(), dirichlet.end());
    dirichlet.erase(std::unique(dirichlet.begin(), dirichlet.end()), dirichlet.end());
    const int nv = M.nv;
    std::unique_ptr<RefElement> RT;
    std::unique_ptr<RefQuad> RQ;
    if (nv == 4) RQ = std::make_unique<RefQuad>(k, equispaced);
    else RT = std::make_unique<RefElement>(k, equispaced);
    const std::vector<double> &nxi = nv == 4 ? RQ->nxi : RT->nxi, &neta = nv == 4 ? RQ->neta : RT->neta;
    SpaceData V;
    const int nc = M.nt(), np = M.np(), n = nv == 4 ? RQ->n : RT->n, de = k - 1;
    const int nint = nv == 4 ? (k - 1) * (k - 1) : (k - 1) * (k - 2) / 2, f0 = nv + nv * de;
    V.nv = nv; V.k = k; V.nloc = n; V.ncells = nc; V.npoints = np;
    const int gs = V.geom_stride();
    V.geom.resize(gs * static_cast<std::size_t>(nc));
    for (int c = 0; c < nc; ++c)
    {
        const int *t = &M.tri[nv * static_cast<std::size_t>(c)];
        double *g = &V.geom[gs * static_cast<std::size_t>(c)];
        if (nv == 4)
        {
            for (int v = 0; v < 4; ++v) { g[2 * v] = M.xy[2 * t[v]]; g[2 * v + 1] = M.xy[2 * t[v] + 1]; }
            continue;
        }
        g[0] = M.xy[2 * t[0]]; g[1] = M.xy[2 * t[0] + 1];
        g[2] = M.xy[2 * t[1]] - M.xy[2 * t[0]]; g[3] = M.xy[2 * t[2]] - M.xy[2 * t[0]];
        g[4] = M.xy[2 * t[1] + 1] - M.xy[2 * t[0] + 1]; g[5] = M.xy[2 * t[2] + 1] - M.xy[2 * t[0] + 1];
        const double d = g[2] * g[5] - g[3] * g[4];
        if (!(d > 0.0)) bad("LagrangeSpace: cell " + std::to_string(c) + " is degenerate or clockwise");
    }
    // edges in first-met order: slot s = nv c + i; an edge's id is the rank of its smallest slot
    const std::size_t ns = nv * static_cast<std::size_t>(nc);
    std::vector<std::pair<int64_t, int>> ks(ns);
    for (int c = 0; c < nc; ++c)
        for (int i = 0; i < nv; ++i)
        {
            int a, b;
            M.edge(c, i, a, b);
            ks[nv * static_cast<std::size_t>(c) + i] = {edge_key(a, b), nv * c + i};
        }
    std::sort(ks.begin(), ks.end());
    std::vector<int> group_of_slot(ns), first_slot;
    for (std::size_t h = 0; h < ns;)
    {
        std::size_t e = h;
        while (e < ns && ks[e].first == ks[h].first) ++e;
        const int g = static_cast<int>(first_slot.size());
        first_slot.push_back(ks[h].second);               // sorted by (key, slot): the first is the smallest slot
        for (std::size_t q = h; q < e; ++q) group_of_slot[ks[q].second] = g;
        h = e;
    }
    const int ne = static_cast<int>(first_slot.size());
    std::vector<int> order(ne), edge_id(ne);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return first_slot[a] < first_slot[b]; });
    for (int r = 0; r < ne; ++r) edge_id[order[r]] = r;
    std::vector<int> cedge(ns);
    std::vector<std::array<int, 2>> edges(ne);
    for (std::size_t s = 0; s < ns; ++s) cedge[s] = edge_id[group_of_slot[s]];
    for (int g = 0; g < ne; ++g)
    {
        int a, b;
        M.edge(first_slot[g] / nv, first_slot[g] % nv, a, b);
        edges[edge_id[g]] = {std::min(a, b), std::max(a, b)};
    }
    V.nedges = ne;
    V.raw_dim = np + ne * de + nc * nint;
    V.cell_raw.resize(static_cast<std::size_t>(nc) * n);
    for (int c = 0; c < nc; ++c)
    {
        int *d = &V.cell_raw[static_cast<std::size_t>(c) * n];
        for (int v = 0; v < nv; ++v) d[v] = M.tri[nv * static_cast<std::size_t>(c) + v];
        for (int i = 0; i < nv; ++i)
        {
            int a, b;
            M.edge(c, i, a, b);
            const bool flip = a > b;
            for (int j = 0; j < de; ++j) d[nv + i * de + j] = np + cedge[nv * static_cast<std::size_t>(c) + i] * de + (flip ? de - 1 - j : j);
        }
        for (int j = 0; j < nint; ++j) d[f0 + j] = np + ne * de + c * nint + j;
    }
    V.raw_xy.assign(2 * static_cast<std::size_t>(V.raw_dim), std::numeric_limits<double>::quiet_NaN());
    std::vector<char> placed(static_cast<std::size_t>(V.raw_dim), 0);
    for (int v = 0; v < np; ++v) { V.raw_xy[2 * v] = M.xy[2 * v]; V.raw_xy[2 * v + 1] = M.xy[2 * v + 1]; placed[v] = 1; }
    for (int c = 0; c < nc; ++c)
    {
        const double *g = &V.geom[gs * static_cast<std::size_t>(c)];
        for (int l = nv; l < n; ++l)
        {
            const int r = V.cell_raw[static_cast<std::size_t>(c) * n + l];
            if (placed[r]) continue;
            const double xi = nxi[l], eta = neta[l];
            if (nv == 4)
            {
                // the bilinear map (lagrange_space_quad.hpp:155-162)
                const double a0 = (1 - xi) * (1 - eta), a1 = xi * (1 - eta), a2 = xi * eta, a3 = (1 - xi) * eta;
                V.raw_xy[2 * r] = a0 * g[0] + a1 * g[2] + a2 * g[4] + a3 * g[6];
                V.raw_xy[2 * r + 1] = a0 * g[1] + a1 * g[3] + a2 * g[5] + a3 * g[7];
            }
            else
            {
                V.raw_xy[2 * r] = g[0] + g[2] * xi + g[3] * eta;
                V.raw_xy[2 * r + 1] = g[1] + g[4] * xi + g[5] * eta;
            }
            placed[r] = 1;
        }
    }
    // Dirichlet elimination over the facets (exterior edges and segment edges, first met)
    std::vector<char> fixed(static_cast<std::size_t>(V.raw_dim), 0), seen(static_cast<std::size_t>(ne), 0);
    for (int c = 0; c < nc; ++c)
        for (int i = 0; i < nv; ++i)
        {
            const int g = cedge[nv * static_cast<std::size_t>(c) + i];
            const bool ext = M.nbr[nv * static_cast<std::size_t>(c) + i] < 0;
            int a, b;
            M.edge(c, i, a, b);
            const int m = M.marker_of(a, b);
            if (!ext && (m < 0 || seen[g])) continue;
            seen[g] = 1;
            const int mark = m >= 0 ? m : 0;
            if (!((all_exterior && ext) || std::binary_search(dirichlet.begin(), dirichlet.end(), mark))) continue;
            fixed[edges[g][0]] = fixed[edges[g][1]] = 1;
            for (int j = 0; j < de; ++j) fixed[np + g * de + j] = 1;
        }
    std::vector<int> raw2ad(static_cast<std::size_t>(V.raw_dim), -1);
    for (int r = 0; r < V.raw_dim; ++r)
        if (!fixed[r]) { raw2ad[r] = static_cast<int>(V.ad2raw.size()); V.ad2raw.push_back(r); }
    V.dim = static_cast<int>(V.ad2raw.size());
    V.cell_dof.resize(V.cell_raw.size());
    for (std::size_t i = 0; i < V.cell_raw.size(); ++i) V.cell_dof[i] = raw2ad[V.cell_raw[i]];
    return V;
}

// --------------------------------------------------------------------------------- pattern
PatternData build_pattern(const SpaceData &V)
{
    const int n = V.dim, nc = V.ncells, nl = V.nloc;
    if (static_cast<int64_t>(nc) * nl * nl >= (int64_t(1) << 31)) bad("pattern: element matrices exceed 2^31 entries");
    PatternData P;
    P.n = n;
    // dof -> (cell, local) in ascending cell order, by counting
    P.vptr.assign(static_cast<std::size_t>(n) + 1, 0);
    for (int c = 0; c < nc; ++c)
        for (int a = 0; a < nl; ++a) { const int i = V.cell_dof[static_cast<std::size_t>(c) * nl + a]; if (i >= 0) ++P.vptr[i + 1]; }
    for (int i = 0; i < n; ++i) P.vptr[i + 1] += P.vptr[i];
    P.vlist.resize(P.vptr[n]);
    {
        std::vector<int> f(P.vptr.begin(), P.vptr.end() - 1);
        for (int c = 0; c < nc; ++c)
            for (int a = 0; a < nl; ++a) { const int i = V.cell_dof[static_cast<std::size_t>(c) * nl + a]; if (i >= 0) P.vlist[f[i]++] = c * nl + a; }
    }
    // per row: (column, element entry) pairs in element order, stably sorted by column
    std::vector<int> rownnz(static_cast<std::size_t>(n)), rowcon(static_cast<std::size_t>(n));
    auto row_pairs = [&](int i, std::vector<std::pair<int, int>> &buf) {
        buf.clear();
        for (int q = P.vptr[i]; q < P.vptr[i + 1]; ++q)
        {
            const int c = P.vlist[q] / nl, a = P.vlist[q] % nl;
            const int *d = &V.cell_dof[static_cast<std::size_t>(c) * nl];
            for (int b = 0; b < nl; ++b)
                if (d[b] >= 0) buf.emplace_back(d[b], (c * nl + a) * nl + b);
        }
        std::stable_sort(buf.begin(), buf.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    };
#pragma omp parallel
    {
        std::vector<std::pair<int, int>> buf;
#pragma omp for schedule(dynamic, 1024)
        for (int i = 0; i < n; ++i)
        {
            row_pairs(i, buf);
            int u = 0;
            for (std::size_t t = 0; t < buf.size(); ++t) if (t == 0 || buf[t].first != buf[t - 1].first) ++u;
            rownnz[i] = u; rowcon[i] = static_cast<int>(buf.size());
        }
    }
    P.indptr.assign(static_cast<std::size_t>(n) + 1, 0);
    std::vector<int64_t> conoff(static_cast<std::size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i) { P.indptr[i + 1] = P.indptr[i] + rownnz[i]; conoff[i + 1] = conoff[i] + rowcon[i]; }
    if (conoff[n] >= (int64_t(1) << 31)) bad("pattern: too many element contributions");
    P.indices.resize(P.indptr[n]);
    P.mptr.resize(static_cast<std::size_t>(P.indptr[n]) + 1);
    P.mlist.resize(conoff[n]);
    P.mptr[P.indptr[n]] = static_cast<int>(conoff[n]);
#pragma omp parallel
    {
        std::vector<std::pair<int, int>> buf;
#pragma omp for schedule(dynamic, 1024)
        for (int i = 0; i < n; ++i)
        {
            row_pairs(i, buf);
            int p = P.indptr[i] - 1;
            int64_t q = conoff[i];
            for (std::size_t t = 0; t < buf.size(); ++t, ++q)
            {
                if (t == 0 || buf[t].first != buf[t - 1].first) { ++p; P.indices[p] = buf[t].first; P.mptr[p] = static_cast<int>(q); }
                P.mlist[q] = buf[t].second;
            }
        }
    }
    return P;
}

// ------------------------------------------------------------------------ boundary facets
FacetData boundary_facets(const MeshData &M, const SpaceData &V, const PatternData &P, const RefElement *RT,
                          const RefQuad *RQ, int degree, std::vector<int> markers, bool all)
{
    std::sort(markers.begin(), markers.end());
    const int nv = M.nv, nl = V.nloc, nc = M.nt();
    FacetData F;
    F.nvar = nv;
    F.nloc = nl;
    std::vector<double> g, gw;
    gauss_legendre(std::max(1, (degree + (nv == 4 ? 3 : 2)) / 2), g, gw);
    F.nq = static_cast<int>(g.size());
    const int nq = F.nq;
    // the reference points of every local edge and the basis there (fe/cache_2d.hpp:135-149)
    static const double tvx[3] = {0.0, 1.0, 0.0}, tvy[3] = {0.0, 0.0, 1.0};
    static const double qvx[4] = {0.0, 1.0, 1.0, 0.0}, qvy[4] = {0.0, 0.0, 1.0, 1.0};
    std::vector<double> fx(static_cast<std::size_t>(nv) * nq), fy(fx.size());
    F.tab.assign(static_cast<std::size_t>(nv) * nq * 3 * nl, 0.0);
    for (int i = 0; i < nv; ++i)
    {
        const int va = nv == 4 ? i : (i + 1) % 3, vb = nv == 4 ? (i + 1) % 4 : (i + 2) % 3;
        const double ax = nv == 4 ? qvx[va] : tvx[va], ay = nv == 4 ? qvy[va] : tvy[va];
        const double bx = nv == 4 ? qvx[vb] : tvx[vb], by = nv == 4 ? qvy[vb] : tvy[vb];
        for (int q = 0; q < nq; ++q)
        {
            const double t = 0.5 * (g[q] + 1.0);
            const std::size_t k = static_cast<std::size_t>(i) * nq + q;
            fx[k] = (1.0 - t) * ax + t * bx;
            fy[k] = (1.0 - t) * ay + t * by;
            double *o = &F.tab[k * 3 * nl];
            if (nv == 4) RQ->tabulate(fx[k], fy[k], o);
            else RT->tabulate(fx[k], fy[k], o);
        }
    }
    for (int c = 0; c < nc; ++c)
        for (int i = 0; i < nv; ++i)
        {
            if (M.nbr[static_cast<std::size_t>(c) * nv + i] >= 0) continue;
            int a = 0, b = 0;
            M.edge(c, i, a, b);
            const int m = M.marker_of(a, b);
            if (!all && !std::binary_search(markers.begin(), markers.end(), m >= 0 ? m : 0)) continue;
            F.cell.push_back(c);
            F.var.push_back(i);
        }
    F.nent = static_cast<int>(F.cell.size());
    const std::size_t np = static_cast<std::size_t>(F.nent) * nq;
    F.x.resize(np); F.y.resize(np); F.w.resize(np); F.nx.resize(np); F.ny.resize(np); F.Ji.resize(4 * np);
    const int gs = V.geom_stride();
    for (int e = 0; e < F.nent; ++e)
    {
        const int c = F.cell[e], i = F.var[e];
        int a = 0, b = 0;
        M.edge(c, i, a, b);
        const double dx = M.xy[2 * b] - M.xy[2 * a], dy = M.xy[2 * b + 1] - M.xy[2 * a + 1], len = std::hypot(dx, dy);
        const double *gc = &V.geom[gs * static_cast<std::size_t>(c)];
        for (int q = 0; q < nq; ++q)
        {
            const std::size_t k = static_cast<std::size_t>(e) * nq + q;
            const double xi = fx[static_cast<std::size_t>(i) * nq + q], eta = fy[static_cast<std::size_t>(i) * nq + q];
            double J[4];
            if (nv == 4)
            {
                // the bilinear map and its Jacobian (fe/lagrange_space_quad.hpp:155-172)
                const double a0 = (1 - xi) * (1 - eta), a1 = xi * (1 - eta), a2 = xi * eta, a3 = (1 - xi) * eta;
                F.x[k] = a0 * gc[0] + a1 * gc[2] + a2 * gc[4] + a3 * gc[6];
                F.y[k] = a0 * gc[1] + a1 * gc[3] + a2 * gc[5] + a3 * gc[7];
                J[0] = (gc[2] - gc[0]) * (1 - eta) + (gc[4] - gc[6]) * eta;
                J[1] = (gc[6] - gc[0]) * (1 - xi) + (gc[4] - gc[2]) * xi;
                J[2] = (gc[3] - gc[1]) * (1 - eta) + (gc[5] - gc[7]) * eta;
                J[3] = (gc[7] - gc[1]) * (1 - xi) + (gc[5] - gc[3]) * xi;
            }
            else
            {
                // the affine map (fe/lagrange_space_2d.hpp:82-89)
                F.x[k] = gc[0] + gc[2] * xi + gc[3] * eta;
                F.y[k] = gc[1] + gc[4] * xi + gc[5] * eta;
                J[0] = gc[2]; J[1] = gc[3]; J[2] = gc[4]; J[3] = gc[5];
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
    // gather lists: per dof the (entry, local) pairs in ascending entry order, then per row the
    // (column, element entry) pairs stably sorted by column
    const int n = V.dim;
    std::vector<int> cnt(static_cast<std::size_t>(n) + 1, 0);
    for (int e = 0; e < F.nent; ++e)
        for (int l = 0; l < nl; ++l) { const int d = V.cell_dof[static_cast<std::size_t>(F.cell[e]) * nl + l]; if (d >= 0) ++cnt[d + 1]; }
    for (int d = 0; d < n; ++d) cnt[d + 1] += cnt[d];
    std::vector<int> lst(cnt[n]);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < F.nent; ++e)
            for (int l = 0; l < nl; ++l)
            {
                const int d = V.cell_dof[static_cast<std::size_t>(F.cell[e]) * nl + l];
                if (d >= 0) lst[fill[d]++] = e * nl + l;
            }
    }
    std::vector<std::pair<int, int>> buf;
    for (int d = 0; d < n; ++d)
    {
        if (cnt[d] == cnt[d + 1]) continue;
        F.vdof.push_back(d);
        F.vptr.push_back(static_cast<int>(F.vlist.size()));
        for (int q = cnt[d]; q < cnt[d + 1]; ++q) F.vlist.push_back(lst[q]);
        buf.clear();
        for (int q = cnt[d]; q < cnt[d + 1]; ++q)
        {
            const int e = lst[q] / nl, l = lst[q] % nl;
            const int *cd = &V.cell_dof[static_cast<std::size_t>(F.cell[e]) * nl];
            for (int j = 0; j < nl; ++j)
                if (cd[j] >= 0) buf.emplace_back(cd[j], (e * nl + l) * nl + j);
        }
        std::stable_sort(buf.begin(), buf.end(), [](const auto &u, const auto &v) { return u.first < v.first; });
        const int *row = P.indices.data() + P.indptr[d], len = P.indptr[d + 1] - P.indptr[d];
        for (std::size_t t = 0; t < buf.size(); ++t)
        {
            if (t == 0 || buf[t].first != buf[t - 1].first)
            {
                const int *hit = std::lower_bound(row, row + len, buf[t].first);
                if (hit == row + len || *hit != buf[t].first) bad("boundary facets: an entry outside the pattern");
                F.mptr.push_back(static_cast<int>(F.mlist.size()));
                F.mpos.push_back(P.indptr[d] + static_cast<int>(hit - row));
            }
            F.mlist.push_back(buf[t].second);
        }
    }
    F.vptr.push_back(static_cast<int>(F.vlist.size()));
    F.mptr.push_back(static_cast<int>(F.mlist.size()));
    return F;
}

std::vector<int> spmv_row_blocks(int n, const int *indptr, int max_rows, int cap)
{
    std::vector<int> b{0};
    int r = 0;
    while (r < n)
    {
        int e = r;
        while (e < n && e - r < max_rows && indptr[e + 1] - indptr[r] <= cap) ++e;
        if (e == r) e = r + 1;                        // a row longer than cap, alone
        b.push_back(e);
        r = e;
    }
    return b;
}

void spmv_tile_columns(const int *indptr, const int *indices, const std::vector<int> &rows, int cap, std::vector<int> &uptr,
                       std::vector<int> &ucol, std::vector<uint16_t> &lidx)
{
    if (cap > 65536) bad("spmv_tile_columns: cap above 65536");
    const int nt = static_cast<int>(rows.size()) - 1;
    const int64_t nnz = nt >= 0 ? indptr[rows.back()] : 0;
    std::vector<int> count(static_cast<std::size_t>(std::max(nt, 0)) + 1, 0);
    lidx.assign(static_cast<std::size_t>(nnz), 0);
    std::vector<std::vector<int>> lists(static_cast<std::size_t>(std::max(nt, 0)));
#pragma omp parallel for schedule(dynamic, 16)
    for (int b = 0; b < nt; ++b)
    {
        const int p0 = indptr[rows[b]], p1 = indptr[rows[b + 1]];
        if (p1 - p0 > cap) continue;
        std::vector<int> &u = lists[b];
        u.assign(indices + p0, indices + p1);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        for (int p = p0; p < p1; ++p)
            lidx[p] = static_cast<uint16_t>(std::lower_bound(u.begin(), u.end(), indices[p]) - u.begin());
        count[b + 1] = static_cast<int>(u.size());
    }
    uptr.assign(count.size(), 0);
    for (int b = 0; b < nt; ++b) uptr[b + 1] = uptr[b] + count[b + 1];
    ucol.resize(static_cast<std::size_t>(nt > 0 ? uptr[nt] : 0));
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nt; ++b)
        std::copy(lists[b].begin(), lists[b].end(), ucol.begin() + uptr[b]);
}

// --------------------------------------------------------------------------- the inner CG's order
bool first_met_order(int n, const int *ip, const int *ix, const std::vector<int64_t> &segments, std::vector<int> &perm,
                     std::vector<int> &ip2, std::vector<int> &ix2, std::vector<int> &emap)
{
    const std::vector<int64_t> seg = segments.empty() ? std::vector<int64_t>{0, int64_t(n)} : segments;
    const int nseg = static_cast<int>(seg.size()) - 1;
    const int nnz = n > 0 ? ip[n] : 0;
    perm.assign(size_t(n), 0);
    std::vector<int> iperm(size_t(n), -1);
    int outside = 0;
    // within each segment: columns in the order the rows first meet them, then the unmet ones
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : outside)
    for (int k = 0; k < nseg; ++k)
    {
        const int r0 = static_cast<int>(seg[size_t(k)]), r1 = static_cast<int>(seg[size_t(k) + 1]);
        int next = r0;
        for (int e = ip[r0]; e < ip[r1]; ++e)
        {
            const int j = ix[e];
            if (j < r0 || j >= r1)
            {
                ++outside;
                break;
            }
            if (iperm[size_t(j)] < 0)
            {
                iperm[size_t(j)] = next;
                perm[size_t(next++)] = j;
            }
        }
        for (int j = r0; j < r1; ++j)
            if (iperm[size_t(j)] < 0)
            {
                iperm[size_t(j)] = next;
                perm[size_t(next++)] = j;
            }
    }
    if (outside) return false;
    ip2.assign(size_t(n) + 1, 0);
    for (int k = 0; k < n; ++k) ip2[size_t(k) + 1] = ip2[size_t(k)] + (ip[perm[size_t(k)] + 1] - ip[perm[size_t(k)]]);
    ix2.assign(size_t(nnz), 0);
    emap.assign(size_t(nnz), 0);
    // row k of P A P^T is row perm[k] with its columns renumbered, in ascending order
#pragma omp parallel
    {
        std::vector<std::pair<int, int>> buf;
#pragma omp for schedule(dynamic, 4096)
        for (int k = 0; k < n; ++k)
        {
            const int b = ip[perm[size_t(k)]], len = ip[perm[size_t(k)] + 1] - b, o = ip2[size_t(k)];
            buf.resize(size_t(len));
            for (int t = 0; t < len; ++t) buf[size_t(t)] = {iperm[size_t(ix[b + t])], b + t};
            std::sort(buf.begin(), buf.end());
            for (int t = 0; t < len; ++t)
            {
                ix2[size_t(o + t)] = buf[size_t(t)].first;
                emap[size_t(o + t)] = buf[size_t(t)].second;
            }
        }
    }
    return true;
}

// --------------------------------------------------------------------------------- orderings
// sparse/amd.hpp:166-234: symmetric adjacency without the diagonal; seeds by increasing degree
// (stable); George-Liu pseudo-peripheral search (at most 32 sweeps); breadth first with the new
// neighbours of a node by increasing degree (stable); the whole order reversed.
std::vector<int> rcm_order(int n, const int *ip, const int *ix)
{
    // adjacency = pattern + transpose, diagonal dropped, sorted unique: counting passes
    std::vector<int> cnt(static_cast<std::size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int p = ip[i]; p < ip[i + 1]; ++p)
        {
            const int j = ix[p];
            if (j < 0 || j >= n) bad("rcm: column index out of range");
            if (j != i) { ++cnt[i + 1]; ++cnt[j + 1]; }
        }
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> adj(cnt[n]), fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int p = ip[i]; p < ip[i + 1]; ++p)
            if (ix[p] != i) { adj[fill[i]++] = ix[p]; adj[fill[ix[p]]++] = i; }
    std::vector<int> aptr(static_cast<std::size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i)
    {
        std::sort(adj.begin() + cnt[i], adj.begin() + cnt[i + 1]);
        const int u = static_cast<int>(std::unique(adj.begin() + cnt[i], adj.begin() + cnt[i + 1]) - (adj.begin() + cnt[i]));
        aptr[i + 1] = aptr[i] + u;
        std::copy(adj.begin() + cnt[i], adj.begin() + cnt[i] + u, adj.begin() + aptr[i]);   // compaction moves left only
    }
    std::vector<int> deg(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) deg[i] = aptr[i + 1] - aptr[i];
    std::vector<int> level(static_cast<std::size_t>(n), -1), queue(static_cast<std::size_t>(n)), order;
    std::vector<char> numbered(static_cast<std::size_t>(n), 0);
    order.reserve(n);
    // breadth-first levels from s over unnumbered nodes: eccentricity, last level in queue[fl, end)
    auto bfs = [&](int s, int &fl, int &end) {
        int head = 0, tail = 0, ecc = 0;
        queue[tail++] = s; level[s] = 0; fl = 0;
        while (head < tail)
        {
            const int v = queue[head++];
            if (level[v] > ecc) { ecc = level[v]; fl = head - 1; }
            for (int p = aptr[v]; p < aptr[v + 1]; ++p)
            {
                const int w = adj[p];
                if (numbered[w] || level[w] >= 0) continue;
                level[w] = level[v] + 1;
                queue[tail++] = w;
            }
        }
        end = tail;
        for (int t = 0; t < tail; ++t) level[queue[t]] = -1;
        return ecc;
    };
    std::vector<int> seeds(static_cast<std::size_t>(n));
    std::iota(seeds.begin(), seeds.end(), 0);
    std::stable_sort(seeds.begin(), seeds.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int> nb;
    for (int seed : seeds)
    {
        if (numbered[seed]) continue;
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
        order.push_back(s); numbered[s] = 1;
        while (head < order.size())
        {
            const int v = order[head++];
            nb.clear();
            for (int p = aptr[v]; p < aptr[v + 1]; ++p)
                if (!numbered[adj[p]]) { nb.push_back(adj[p]); numbered[adj[p]] = 1; }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

Permuted permute_pattern(int n, const int *ip, const int *ix, const std::vector<int> &p)
{
    std::vector<int> pinv(static_cast<std::size_t>(n), -1);
    for (int k = 0; k < n; ++k)
    {
        if (p[k] < 0 || p[k] >= n || pinv[p[k]] >= 0) bad("permuted: not a permutation");
        pinv[p[k]] = k;
    }
    Permuted B;
    B.indptr.assign(static_cast<std::size_t>(n) + 1, 0);
    for (int k = 0; k < n; ++k) B.indptr[k + 1] = B.indptr[k] + (ip[p[k] + 1] - ip[p[k]]);
    B.indices.resize(B.indptr[n]); B.src.resize(B.indptr[n]);
    std::vector<std::pair<int, int>> row;
    for (int k = 0; k < n; ++k)
    {
        row.clear();
        for (int q = ip[p[k]]; q < ip[p[k] + 1]; ++q) row.emplace_back(pinv[ix[q]], q);
        std::sort(row.begin(), row.end());
        for (std::size_t t = 0; t < row.size(); ++t) { B.indices[B.indptr[k] + t] = row[t].first; B.src[B.indptr[k] + t] = row[t].second; }
    }
    return B;
}

namespace {
Levels bucket(const std::vector<int> &lev, int nlev)
{
    Levels L;
    const int n = static_cast<int>(lev.size());
    L.ptr.assign(static_cast<std::size_t>(nlev) + 1, 0);
    for (int i = 0; i < n; ++i) ++L.ptr[lev[i] + 1];
    for (int l = 0; l < nlev; ++l) L.ptr[l + 1] += L.ptr[l];
    L.rows.resize(n);
    std::vector<int> f(L.ptr.begin(), L.ptr.end() - 1);
    for (int i = 0; i < n; ++i) L.rows[f[lev[i]]++] = i;
    return L;
}
} // namespace

Levels lower_levels(int n, const int *ip, const int *ix)
{
    std::vector<int> lev(static_cast<std::size_t>(n), 0);
    int nlev = n > 0 ? 1 : 0;
    for (int i = 0; i < n; ++i)
    {
        int l = 0;
        for (int p = ip[i]; p < ip[i + 1] && ix[p] < i; ++p) l = std::max(l, lev[ix[p]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    return bucket(lev, nlev);
}

Levels upper_levels(int n, const int *ip, const int *ix)
{
    std::vector<int> lev(static_cast<std::size_t>(n), 0);
    int nlev = n > 0 ? 1 : 0;
    for (int i = n - 1; i >= 0; --i)
    {
        int l = 0;
        for (int p = ip[i + 1] - 1; p >= ip[i] && ix[p] > i; --p) l = std::max(l, lev[ix[p]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    return bucket(lev, nlev);
}

} // namespace femd_amd
