// host.hpp -- the host-side data of the MI355X path: meshes, the P_k numbering, the reference
// element tables, quadrature rules, the CSR pattern with the element-order gather lists the
// assembly kernels read, the reverse Cuthill-McKee order and the level schedules of the
// triangular sweeps.  Everything here runs once per mesh, space or matrix (the reference
// caches the same things: pattern per space pair, scatter map per assembler).
#pragma once
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace femd_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// A triangulation in the state mesh_from_arrays leaves it (mesh/structured.hpp:99-221):
// counter-clockwise triangles, neighbour across the edge opposite each vertex, segments with
// markers (user segments first, then the remaining boundary edges in loop order).  Or a
// quadrilateral mesh as QuadMesh holds it (mesh/quad_mesh.hpp:4-103): counter-clockwise convex
// quads, neighbour across edge i (from vertex i to vertex i+1), the listed segments first, then
// the remaining boundary edges in quad order.
struct MeshData {
    int nv = 3;                      // vertices per cell
    std::vector<double> xy;          // np x 2
    std::vector<int> tri, nbr;       // nt x nv
    std::vector<int> seg, smark;     // ns x 2, ns
    std::vector<int64_t> seg_key;    // sorted edge keys of the segments
    std::vector<int> seg_of_key;     // segment index for seg_key[i]
    int np() const { return static_cast<int>(xy.size() / 2); }
    int nt() const { return static_cast<int>(tri.size() / nv); }
    int marker_of(int a, int b) const;        // the segment's marker, -1 when (a, b) is no segment
    // the vertices of local edge i of cell c
    void edge(int c, int i, int &a, int &b) const
    {
        const int *t = &tri[static_cast<std::size_t>(c) * nv];
        if (nv == 3) { a = t[(i + 1) % 3]; b = t[(i + 2) % 3]; }
        else { a = t[i]; b = t[(i + 1) % 4]; }
    }
};

MeshData mesh_from_arrays(std::vector<double> xy, std::vector<int> tri, const std::vector<int> &segs,
                          const std::vector<int> &smarks, int marker);
MeshData mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, int diag, const int markers[4]);
MeshData quad_mesh_from_arrays(std::vector<double> xy, std::vector<int> quads, const std::vector<int> &segs,
                               const std::vector<int> &smarks, int marker);
// python/femd/quadmesh.py:146-170: rectangle_mesh's grid and sides, quad (i, j) = j nx + i
MeshData quad_mesh_rectangle(double x0, double y0, double x1, double y1, int nx, int ny, const int markers[4]);

// The nodal P_k basis on the reference triangle (fe/simplex_basis.hpp: Dubiner modes times the
// inverse Vandermonde matrix; equispaced nodes, or symmetrized Gauss-Lobatto edge nodes and
// warp-and-blend interior nodes; local order vertices / edges / interior).
struct RefElement {
    int k = 1, n = 3;
    std::vector<double> nxi, neta, vinv;
    explicit RefElement(int degree, bool equispaced = true);
    void tabulate(double xi, double eta, double *out3n) const;     // values, d/dxi, d/deta
};

// The nodal Q_k basis on the reference square [0,1]^2 (fe/lagrange_space_quad.hpp:28-110): 1D
// Lagrange polynomials in barycentric form on k+1 symmetrized equispaced or Gauss-Lobatto points,
// tensor products;
// local order vertices, the nodes of each edge along it, interior row by row.
struct RefQuad {
    int k = 1, n = 4;
    std::vector<double> t, w, nxi, neta;
    std::vector<int> ii, jj;
    explicit RefQuad(int degree, bool equispaced = true);
    void tabulate(double xi, double eta, double *out3n) const;     // values, d/dxi, d/deta
};

struct QuadRule {
    std::vector<double> xi, eta, w;
    int size() const { return static_cast<int>(w.size()); }
};
QuadRule triangle_rule(int degree);
QuadRule square_rule(int degree);            // fe/cache_2d.hpp:75-84, tensor Gauss-Legendre

// A continuous P_k (triangles) or Q_k (quadrilaterals) space (fe/space_2d.hpp:390-526,
// fe/lagrange_space_2d.hpp:32-67, fe/lagrange_space_quad.hpp:112-210).
struct SpaceData {
    int nv = 3, k = 1, nloc = 3, ncells = 0, npoints = 0, nedges = 0, raw_dim = 0, dim = 0;
    std::vector<int> cell_raw, cell_dof, ad2raw;
    std::vector<double> raw_xy;
    // per cell: triangles x0, y0, J00, J01, J10, J11 (6 doubles); quadrilaterals the four
    // vertices x0, y0, .., x3, y3 (8 doubles; J varies in the cell)
    std::vector<double> geom;
    int geom_stride() const { return nv == 4 ? 8 : 6; }
};
SpaceData build_space(const MeshData &M, int k, std::vector<int> dirichlet, bool all_exterior, bool equispaced = true);

// The pattern of a space with itself and, for every stored entry, the element-matrix entries
// that add into it in ascending element order (the order assemble_matrix_2d adds them,
// forms/assembly_2d.hpp:265-301): entry p receives E[mlist[q]] for q in [mptr[p], mptr[p+1]),
// E the element matrices laid out [e][i][j].  The same for vectors: dof i receives
// F[vlist[q]], q in [vptr[i], vptr[i+1]), F laid out [e][i].
struct PatternData {
    int n = 0;
    std::vector<int> indptr, indices;
    std::vector<int> mptr, mlist;
    std::vector<int> vptr, vlist;
    int64_t nnz() const { return indptr.empty() ? 0 : indptr.back(); }
};
PatternData build_pattern(const SpaceData &V);

// The exterior facets of a space whose side marker is in `markers` (every exterior facet with
// `all`), in the order of the space's facet list (fe/space_2d.hpp:432-448: cells ascending, local
// edges ascending), with Gauss-Legendre points along each edge and the geometry of
// fe/cache_2d.hpp:126-183 at every point: physical point, weight (reference weight times half the
// edge length), outward unit normal and J^{-1} of the cell's map; the basis tabulated once per
// local edge; and the gather lists that add the facet element matrices (vectors) into a pattern's
// stored entries (dofs) in ascending facet order.
struct FacetData {
    int nq = 0, nent = 0, nvar = 0, nloc = 0;
    std::vector<int> cell, var;                  // entry e: its cell and local edge
    std::vector<double> x, y, w, nx, ny, Ji;     // per (entry, point), [e nq + q]; Ji 4 per point
    std::vector<double> tab;                     // [(variant nq + q) 3 nloc]: values, d/dxi, d/deta
    // stored entry mpos[k] receives E[mlist[j]], j in [mptr[k], mptr[k+1]), E laid out [e][i][j];
    // dof vdof[k] receives F[vlist[j]], j in [vptr[k], vptr[k+1]), F laid out [e][i]
    std::vector<int> mpos, mptr, mlist, vdof, vptr, vlist;
};
FacetData boundary_facets(const MeshData &M, const SpaceData &V, const PatternData &P, const RefElement *RT,
                          const RefQuad *RQ, int degree, std::vector<int> markers, bool all);

// Row blocks of the CSR-stream SpMV: block b covers rows [rows[b], rows[b+1]) whose entries
// number at most cap; a single row longer than cap gets a block of its own (flagged long).
std::vector<int> spmv_row_blocks(int n, const int *indptr, int max_rows, int cap);

// The distinct columns of each tile (rows [rows[b], rows[b+1])): ucol[uptr[b] .. uptr[b+1]) sorted
// ascending, and for every entry p of the tile lidx[p] = the position of indices[p] in that list.
// Tiles with more than cap entries (a single long row) get no list and lidx 0.
void spmv_tile_columns(const int *indptr, const int *indices, const std::vector<int> &rows, int cap, std::vector<int> &uptr,
                       std::vector<int> &ucol, std::vector<uint16_t> &lidx);

// sparse/amd.hpp:166-234 semantics: p[k] is the original row that becomes row k.
std::vector<int> rcm_order(int n, const int *indptr, const int *indices);
// The inner CG's order of a square pattern's unknowns: within every segment (row ranges; none: one),
// the order in which a row-by-row scan first meets them as columns, those it never meets last in
// their own order (perm: new -> old), and the reordered pattern P A P^T (ip2, ix2; each row's
// columns ascending) with emap[p] the entry of the original behind entry p of the reordered one.
// false when a segment's rows have columns outside it.
bool first_met_order(int n, const int *ip, const int *ix, const std::vector<int64_t> &segments, std::vector<int> &perm,
                     std::vector<int> &ip2, std::vector<int> &ix2, std::vector<int> &emap);

// B = P A P^T on the pattern level (csr.hpp:286-313): the new pattern and, per new entry,
// the position of its value in A.
struct Permuted {
    std::vector<int> indptr, indices, src;
};
Permuted permute_pattern(int n, const int *indptr, const int *indices, const std::vector<int> &p);

// Level sets of the lower (row i waits for the columns j < i of its row) or the upper
// (columns j > i) triangular sweep: rows listed level by level, ascending within a level.
struct Levels {
    std::vector<int> ptr, rows;
    int count() const { return static_cast<int>(ptr.size()) - 1; }
};
Levels lower_levels(int n, const int *indptr, const int *indices);
Levels upper_levels(int n, const int *indptr, const int *indices);

} // namespace femd_amd
