// cg_engine.hpp -- the inner CG of the mixed-precision solves, the batched solver and the time
// stepper: Jacobi-preconditioned CG on the symmetrically scaled system
//     (S A S) y = S r,  S = diag(a_ii)^(-1/2),  d = S y
// (the iterates of PCG with M = diag(A) in exact arithmetic), in FP32 or FP64, for one matrix or
// for many independent systems at once (the systems of a block-diagonal matrix whose pattern has
// one segment per system; every system keeps its own scalars and state).
//
// Iteration of every running system (two or three kernels, reductions in a fixed order):
//   product  p' = r + beta p at the gathered columns and the own rows (into the other direction
//            buffer), q = (S A S) p', partials of p'.q per tile
//   alpha    per system: p'.q from its tiles' partials, alpha = r.r / p'.q (for small systems,
//            and one large one, every update workgroup forms it instead)
//   update   y += alpha p', r -= alpha q, partials of r.r and of r.D r per chunk (D = diag(A));
//            the last of a system's workgroups to finish (counted in by device-scope atomics)
//            sums them: pause when r.r has dropped by inner_tol^2 since the segment began, or
//            when rho^2 r.D r -- the squared norm of the unscaled residual b - A x the recurrence
//            stands for -- has reached the solve's target (tol ||b||)^2, or at the iteration cap,
//            or on breakdown; beta = r.r_new / r.r
// A paused system's kernels return at once, so a fixed launch sequence (and a captured graph of
// it) serves any mix of running and finished systems.
//
// Between segments the solution is corrected in FP64 (x += rho S y), the FP64 residual b - A x
// recomputed and the system either retired or resumed from it (reliable updates): r = S r_64 /
// rho_new, y = 0, the direction kept (scaled by rho / rho_new) with the beta that makes the next
// direction A-conjugate to it, -(r.q)/(p.q).  A breakdown restarts the system from p = 0.
//
// The inner vectors (r, p, q, y), the scaled values and diag(A) are kept in the pattern's
// LocalOrder (every system's unknowns in the order its rows first meet them as columns) unless
// the caller or femd_inner_order asks for the original order; the FP64 vectors (b, x, the
// residual) stay in the caller's order and meet the inner ones only in the rescale and fold-back
// kernels.
#pragma once
#include "internal.hpp"

namespace femd_amd {

// The state of one system.  reason: 0 running, 1 r.r dropped by inner_tol^2 or the residual
// estimate met its target (a reliable update is due), 2 breakdown, 3 the segment's iteration cap.
struct CGSys {
    double rr;           // r.r of the current residual (scaled system)
    double alpha, beta;  // alpha of the current iteration; beta for the next product
    double thr;          // pause once r.r <= thr
    double tgt;          // or once r.D r <= tgt  ((tol ||b|| / rho)^2; 0: no such test)
    double tol2;         // inner_tol^2
    double rho;          // the FP64 residual norm the system is scaled by
    int it, total, max_iter, active, reason;   // it: since the last restart; total: all
    int dir;             // the direction buffer (P0 / P1) holding the current direction
    int pad[2];
};

// S = diag(A)^(-1/2) (FP64) and S A S (FP32 / FP64) on the matrix, in the pattern's LocalOrder when
// `reorder` (and femd_inner_order allows it), else in place; cached until the values or the order
// change; FEMD_ENOTSPD when a diagonal entry is not positive.
void scaled_values(femd_csr_s *A, hipStream_t s, bool reorder = true);
void scaled_values64(femd_csr_s *A, hipStream_t s, bool reorder = true);

struct CGEngine {
    struct Result {
        int converged = 0, iterations = 0, updates = 0, rounds = 0;
        double worst = 0.0;
    };

    femd_csr_s *A = nullptr;
    int precision = FEMD_FP32;
    int np = 0;
    int64_t n = 0, max_rows = 0;
    std::vector<int64_t> roff;          // system row offsets (np + 1)
    int nch = 0;                        // vector chunks (1024 rows at most, one system each)
    std::vector<int> h_cptr;
    DevBuf clo, cpid, cptr;
    DevPattern *pat = nullptr;          // A's pattern in the inner CG's order (its LocalOrder)
    const int *perm = nullptr;          // position -> unknown of A (device)
    const SpmvPlan *plan = nullptr;     // pat's plan
    uint64_t plan_id = 0;
    DevBuf seg_tiles;
    DevBuf st, tparts, cparts, r, q, P0, P1, y, r64, t64, sel, retired, rho_new, tgt, cnt;
    bool use_dict = false;              // the product reads the scaled values through their dictionary
    bool dict_raw = false;              // ... and some tiles, with too many values, read the values
    bool fuse_alpha = false;            // small systems: every update workgroup forms alpha itself
    int update_per = 1;                 // chunks per update workgroup (> 1: one large system, alpha fused)
    int parity = 0;                     // the next iteration reads its direction from P1 (1) or P0 (0)
    DevBuf list, aoff;                  // the product's tile list (two int4 per tile) and its offsets
    DevBuf acnt, abase, anw;            // the update's arrival counters, each system's first and its workgroups
    int list_count = -1;
    int dict_gen = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;

    CGEngine() = default;
    CGEngine(const CGEngine &) = delete;
    CGEngine &operator=(const CGEngine &) = delete;
    ~CGEngine();

    // Bind to A (square; its pattern's segments are the systems, none means one system) in a
    // precision, working in the pattern's LocalOrder when `reorder` (and femd_inner_order allows
    // it); allocates the work space.  Re-binding to the same matrix keeps it.
    void setup(femd_csr_s *A, int precision, hipStream_t s, bool reorder = true);
    // count iterations of every running system (graph capturable)
    void iterate(int count, hipStream_t s);
    // GRAPH_ITERS iterations, from a captured graph when use_graph
    void replay(bool use_graph, hipStream_t s);
    // Solve A x = b (every system) to ||b_p - A_p x_p|| <= tol ||b_p|| with reliable updates each
    // time a system's r.r has dropped by inner_tol^2; x holds the initial guess.  max_updates caps
    // the updates per system.  iters / res (per system) may be NULL.
    Result solve(const double *b, double *x, double tol, double inner_tol, int max_updates, bool use_graph, int *iters,
                 double *res, hipStream_t s);

    // The same solve in steps, for a host loop that drives several engines at once: begin (b and x
    // stay in use until the solve has finished), then launch (queues one graph replay and the
    // read-back of the running / paused counts, returns at once) and complete (waits for it, runs
    // the reliable updates of the paused systems and returns the systems retired by this step)
    // until finished().  A retired system's x and system_result are final.
    void begin(const double *b, double *x, double tol, double inner_tol, int max_updates, bool use_graph, hipStream_t s);
    void launch(hipStream_t s);
    std::vector<int> complete(hipStream_t s);
    bool finished() const { return run_.running == 0 && !run_.launched; }
    void system_result(int p, int *iters, double *res, int *converged) const;
    Result result(int *iters, double *res, hipStream_t s);

    // Device-driven pieces for a captured loop (the time stepper), on the systems with
    // sel[sys] != 0: start a segment of at most max_iter iterations from the FP64 residual r64_dev
    // (norms in rho_dev; 0 = nothing to do) -- with no direction (sel 2) or keeping the current one
    // (sel 1, a reliable update) -- pausing at a drop by inner_tol or when the unscaled residual
    // estimate reaches sqrt(tgt_dev[sys]) rho (tgt_dev may be NULL); fold the solution into x.
    void start_device(const double *r64_dev, const double *rho_dev, const double *tgt_dev, const int *sel_dev,
                      double inner_tol, int max_iter, hipStream_t s);
    void finish_device(const int *sel_dev, double *x, hipStream_t s);

  private:
    struct Run {
        const double *b = nullptr;
        double *x = nullptr;
        double tol = 0.0, inner_tol = 0.5;
        int max_updates = 1, max_iter = 1, running = 0, rounds = 0;
        bool use_graph = true, launched = false;
        std::vector<int> sel, retired, conv, updates;
        std::vector<double> rel, nb2;
        std::vector<CGSys> hs;
    } run_;
    std::unique_ptr<Pinned> counts;
    void put_int(DevBuf &d, const std::vector<int> &v, hipStream_t s);
    void bind(femd_csr_s *A, int precision, hipStream_t s);
    void build_list(const std::vector<int> *leave_out, hipStream_t s);
    std::vector<double> norms2(const double *b, const double *x, hipStream_t s);
    void resume(const double *r64_dev, const double *rho_dev, const double *tgt_dev, const int *sel_dev, double tol2,
                int max_iter, hipStream_t s);
    void drop_graph();
};

} // namespace femd_amd
