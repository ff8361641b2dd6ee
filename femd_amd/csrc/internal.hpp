// internal.hpp -- the runtime shared by the C ABI translation units (capi.hip, batch.hip): error
// handling, the device and its caching memory pool, device buffers, and the handle types.
#pragma once
#include "femd_hip.h"
#include "host.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace femd_amd {

extern thread_local std::string g_err;

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw Error(FEMD_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) ::femd_amd::hip_check((x), #x)

[[noreturn]] inline void fail(int code, const std::string &m) { throw Error(code, m); }

// Every entry point runs its body through api(): exceptions become status codes and the message.
template <class F> inline int api(F &&f)
{
    try
    {
        g_err.clear();
        f();
        return FEMD_OK;
    }
    catch (const Error &e) { g_err = e.what(); return e.code; }
    catch (const std::bad_alloc &) { g_err = "out of host memory"; return FEMD_ENOMEM; }
    catch (const std::exception &e) { g_err = e.what(); return FEMD_EINVAL; }
}

// ---------------------------------------------------------------------------- device and pool
struct Device {
    bool ready = false;
    int id = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::multimap<size_t, void *> free_blocks;       // the caching pool: size -> blocks
    size_t cached = 0;
};
extern Device g_dev;
extern std::mutex g_init_mu;

void ensure_device();
inline size_t round_size(size_t b) { return b <= 256 ? 256 : (b + 4095) & ~size_t(4095); }
void *pool_alloc(size_t bytes);
void pool_free(void *p, size_t bytes);

inline hipStream_t S(femd_stream s)
{
    ensure_device();
    return s ? static_cast<hipStream_t>(s) : g_dev.stream;
}

// Device buffer owned by the pool.  A released block is reused only by work queued after the
// release on the same stream; users of other streams synchronize before releasing.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) : p(b ? pool_alloc(b) : nullptr), bytes(b) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() { if (p) pool_free(p, bytes); p = nullptr; bytes = 0; }
    void ensure(size_t b) { if (bytes < b) { release(); p = pool_alloc(b); bytes = b; } }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

template <class T> inline DevBuf upload(const T *h, size_t n, hipStream_t s)
{
    DevBuf d(n * sizeof(T));
    if (n) HIPC(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}
template <class T> inline DevBuf upload(const std::vector<T> &v, hipStream_t s) { return upload(v.data(), v.size(), s); }

// Pinned host scratch for small read-backs.
struct Pinned {
    void *p = nullptr;
    explicit Pinned(size_t b) { HIPC(hipHostMalloc(&p, b, hipHostMallocDefault)); }
    ~Pinned() { if (p) (void)hipHostFree(p); }
    Pinned(const Pinned &) = delete;
    Pinned &operator=(const Pinned &) = delete;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

inline double read_scalar(const double *d, hipStream_t s)
{
    double h = 0.0;
    HIPC(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    return h;
}

// One SpMV launch shape ready to launch on one pattern (spmv.hip): its tile table, the tiles'
// distinct-column lists, and the time the autotuner measured.
struct SpmvPlan {
    uint64_t id = 0;               // unique per built plan (a holder of a plan's buffers can tell a new one)
    int variant = 0;
    int ntiles = 0;
    DevBuf blocks;                 // int2 per tile boundary: (first row, first entry)
    DevBuf uptr, ucol, lidx;       // each tile's distinct columns, and every entry's position in them
    DevBuf tile_sys;               // the segment (system) of every tile, when the pattern has segments
    std::vector<int> seg_tiles;    // tiles of segment k: [seg_tiles[k], seg_tiles[k + 1])
    double ms = 0.0;
};

// Lossless per-tile dictionary of one value array under one tile plan: tile b's distinct values
// (at most 256) in dval[doff[b] .. doff[b + 1]) and every entry's 1-byte position in them (vidx),
// 3 bytes per entry with the column positions instead of 6 (FP32) or 10 (FP64).
struct ValueDict {
    bool valid = false;            // built for the current values
    bool usable = false;           // tiles of more than 256 distinct values hold at most a quarter of the entries
    int gen = 0;                   // bumped by every build
    int max_distinct = 0;          // most distinct values in one tile (385: more than 384, not counted on)
    int raw_tiles = 0;             // tiles with more than 256, which read the value array
    int64_t distinct = 0;          // dictionary entries over all tiles
    uint64_t plan_id = 0;          // the tiling it was built for (SpmvPlan::id)
    DevBuf doff, dval, vidx;
};
// Build d for values v (FP32 or FP64 by `precision`) under plan p (spmv.hip).
void build_dict(const SpmvPlan &p, const void *v, int precision, int64_t nnz, ValueDict &d, hipStream_t s);
// Whether the solvers use value dictionaries where a matrix allows them (femd_value_dictionary).
bool dict_enabled();
// Whether the inner CG works in its own order of the unknowns (femd_inner_order).
bool inner_order_enabled();

struct DevPattern;
// The inner CG's order of a square pattern's unknowns (cg_engine.hip): within every segment, the
// order in which a row-by-row scan of the segment first meets them as columns, those it never meets
// last in their own order.  Consecutive rows then share most of their columns, so a tile of the
// reordered pattern gathers fewer distinct columns, from fewer cache lines, than one of a pattern
// numbered vertices first, then edges, then cell interiors (a P3 tile of 2048 entries: 229 instead
// of 478 columns, 18 instead of 35 lines of 128 bytes).
struct LocalOrder {
    DevBuf perm;                        // new position -> old (n)
    DevBuf emap;                        // entry of the reordered pattern -> entry of the original (nnz)
    std::shared_ptr<DevPattern> pat;    // P A P^T's pattern, each row's columns ascending, same segments
};

struct DevPattern {
    int n = 0, m = 0;
    int64_t nnz = 0;
    std::vector<int64_t> segments;     // row ranges no SpMV tile may straddle (the batch's systems)
    std::vector<int> h_ip, h_ix;
    DevBuf ip, ix, diag;
    bool have_diag = false;
    DevBuf tpos;                       // every entry's transpose position (or -1), for the asymmetry
    bool have_tpos = false;
    std::unique_ptr<SpmvPlan> plan64, plan32;    // tuned launch shapes per precision
    int forced[2] = {-1, -1};                    // shape set by femd_spmv_select (FP64, FP32); -1: tuned
    std::unique_ptr<LocalOrder> local;           // built by the inner CG on first use
};

// The tuned SpMV of a pattern (values give the tuning runs realistic data).  Every shape sums each
// row left to right in the reference's order.
const SpmvPlan &spmv_plan(DevPattern &P, int precision, const void *values, hipStream_t s);

// The inner CG's set-up on the device (setup_dev.hip), on unless femd_device_setup(0).  The first-met
// order of a square pattern's unknowns into L (its pattern's host row pointers included; false
// when a row meets a column outside its segment), and the column lists of tiles
// (blocks: first row, first entry of every tile boundary) -- both equal to host.cpp's.
bool device_setup_enabled();
bool device_first_met_order(int n, const int *ip, const int *ix, int64_t nnz, const std::vector<int64_t> &segments,
                            LocalOrder &L, hipStream_t s);
void device_tile_columns(const int *ix, int64_t nnz, int ntiles, const int2 *blocks, int cap, DevBuf &uptr, DevBuf &ucol,
                         DevBuf &lidx, hipStream_t s);
// A space's pattern with its gather lists (PatternData's arrays, equal to build_pattern's) from its
// cell_dof on the device; the row pointers and columns are also copied to the host.
struct DevPatternLists {
    int n = 0;
    int64_t nnz = 0, ncon = 0;
    DevBuf ip, ix, mptr, mlist, vptr, vlist;
    std::vector<int> h_ip, h_ix;
};
void device_build_pattern(int n, int ncells, int nl, const int *cell_dof_dev, DevPatternLists &out, hipStream_t s);
void spmv_run64(const SpmvPlan &p, const DevPattern &P, const double *v, const double *x, double *y, const int *stop,
                hipStream_t s);
void spmv_run32(const SpmvPlan &p, const DevPattern &P, const float *v, const float *x, float *y, const int *stop,
                hipStream_t s);
// The FP64 product on the tiles whose segment k has mask[k] != 0 only (segmented patterns).
void spmv_run64_masked(const SpmvPlan &p, const DevPattern &P, const double *v, const double *x, double *y, const int *mask,
                       hipStream_t s);
// threads and LDS entries of a plan's launch shape
void plan_shape(const SpmvPlan &p, int *threads, int *cap);
const char *spmv_variant_name(const SpmvPlan &p);
extern const int NSHAPES_PUBLIC;
const char *shape_name(int i);
int shape_index(const char *name, int precision);     // -1: unknown, or not admissible in precision
SpmvPlan shape_plan(const DevPattern &P, int si, int precision, hipStream_t s);

std::shared_ptr<DevPattern> make_pattern(int n, int m, std::vector<int> ip, std::vector<int> ix, hipStream_t s);
const int *diag_of(DevPattern &P, hipStream_t s);

struct Quad {
    int nq = 0;
    std::vector<double> xi, eta, w;
    DevBuf tab, qw, qxi, qeta;
};

// The matrix-core form of one dx group's element matrices (mfma.hip): its combos of reference
// rows, the table R and the output columns, built once per (rule, codes, symmetry).
struct MfmaPlan {
    int ncombo = 0, spc = 0, ks = 0, ntile = 0, nout = 0, nt = 0;
    unsigned alpha = 0, beta = 0;
    bool rlds = false;
    DevBuf rt, ocol;
};

} // namespace femd_amd

// ------------------------------------------------------------------------------ the handles
struct femd_mesh_s {
    femd_amd::MeshData M;
};

struct femd_space_s {
    femd_amd::MeshData M;                          // the mesh it was built on (its boundary facets)
    femd_amd::SpaceData V;
    femd_amd::DevBuf geom;
    femd_amd::DevBuf cell_dof_dev;                 // cell_dof on the device (known fields on this space)
    femd_amd::PatternData P;
    bool have_host_pattern = false, have_pattern = false;
    std::shared_ptr<femd_amd::DevPattern> pat;
    femd_amd::DevBuf mptr, mlist, vptr, vlist;
    std::map<int, std::unique_ptr<femd_amd::Quad>> quads;
    std::map<std::string, std::unique_ptr<femd_amd::MfmaPlan>> mfma_plans;
    std::unique_ptr<femd_amd::RefElement> ref;     // triangles
    std::unique_ptr<femd_amd::RefQuad> refq;       // quadrilaterals
};

struct femd_facets_s {
    femd_space_s *space = nullptr;
    femd_amd::FacetData F;
    femd_amd::DevBuf var, w, Ji, tab, x, y, nx, ny, mpos, mptr, mlist, vdof, vptr, vlist;
};

struct femd_csr_s {
    std::shared_ptr<femd_amd::DevPattern> pat;
    femd_amd::DevBuf v, v32;
    bool v32_valid = false;
    // The inner CG's copies, in the pattern's LocalOrder (entries of the reordered pattern, rows
    // in its order): S A S in FP32, S = diag(A)^(-1/2) (FP64, original order), diag(A) in FP32 (the
    // inner CG's estimate of the true residual), S A S in FP64 (the FP64 inner CG), and the
    // dictionaries of the scaled values under the reordered pattern's plans.
    femd_amd::DevBuf v32s, scale;
    femd_amd::DevBuf diagf;
    bool v32s_valid = false;
    femd_amd::DevBuf v64s;
    bool v64s_valid = false;
    bool scaled_reordered = false;       // v32s, diagf (and v64s) are in the LocalOrder (else in place)
    femd_amd::ValueDict d32s, d64s;
    const femd_space_s *space = nullptr;
};

struct femd_precond_s {
    int kind = 0, ordering = 0, n = 0;
    double omega = 1.0, min_pivot = 0.0;
    femd_amd::DevBuf inv;                               // Jacobi
    femd_amd::DevBuf perm, ip, ix, diag, lu;            // ILU(0) / SSOR on P A P^T
    femd_amd::DevBuf lptrL, lrowsL, lptrU, lrowsU;
    int nlevL = 0, nlevU = 0;
    femd_amd::DevBuf rp, zp;                            // permuted work vectors
    // the planned one-workgroup sweeps (sweep.hpp); when planning fails, the level-set kernels above
    bool sweep = false;
    int nstepsF = 0, nstepsB = 0, nlevF = 0, nlevB = 0;
    uint32_t zunitF = 0, zunitB = 0;     // the streams' zeros
    int eregF = 24, eregB = 24;          // entries per row in registers
    int64_t bytesF = 0, bytesB = 0, entriesLU = 0;
    femd_amd::DevBuf streamF, streamB, firstF, firstB, inposF;
    femd_amd::DevBuf scatF, scatB, swork;    // where the sweeps' results go; their staging vector
    std::string sweep_note;                             // why the sweep was not planned, when it was not
};

namespace femd_amd {
inline femd_csr_s *check_csr(femd_csr A)
{
    if (!A) fail(FEMD_EINVAL, "null matrix");
    return A;
}
const float *values32(femd_csr_s *A, hipStream_t s);
} // namespace femd_amd
