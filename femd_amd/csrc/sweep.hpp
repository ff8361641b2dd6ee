// sweep.hpp -- the triangular sweeps of ILU(0) and SSOR (sparse/precond.hpp:125-144, 189-205) as one
// workgroup's program: the host plans it once per preconditioner (sweep.cpp), the GPU runs it
// (sweep.hip).
//
// A sweep's rows depend on each other only through the entries of their row in the sweep's
// triangle, so the rows of one dependency level can run together; a level is one or more steps of
// at most SW_ROWS rows, separated by workgroup barriers.  Every row sums its products left to right
// in stored order, as the reference does, so the result is the reference's bit for bit whatever
// the schedule.
//
// The z values live in LDS slots: a row's slot is taken at its step and given back after the last
// step that reads it (on C1's RCM order at most ~5.7k values are live at once).  Everything else a
// step needs -- its rows' inputs, destinations, slots, counts, diagonals, entry values and the
// slots of their columns -- comes straight from HBM into registers: the workgroup's waves form
// SW_TEAMS teams of SW_TWAVES, team s % SW_TEAMS runs step s, and each wave loads the rows of its
// team's next step during the step after its own, SW_TEAMS - 1 steps ahead of use.  A step then
// costs one round of LDS gathers of z values, the products subtracted in stored order, and the
// stores of its results.
//
// The stream holds one block per step and wave of the step's team (steps in order, waves in
// order), 16-byte units, for the wave's L lanes (its rows, longest first, L a multiple of 8, 0 for
// a wave with no rows):
//   u32 next[4]     the block of the same wave in its team's next step: unit offset, emax | rows
//                   << 16 (rows: the lanes that hold one), L, and where its results go: its first
//                   row's position in the sweep's order
//   rows[L]         per lane one unit: f64 vin, the row's input (r forward, filled per
//                   application; y backward, filled from the forward sweep's results); u32 0;
//                   u32 sc, the row's own LDS slot (SW_TRASH when no later row reads it) | entries
//                   << 16
//   f64 diag[L]     the row's diagonal (kinds that divide by it; L / 2 units)
//   pairs           entries 2q and 2q + 1 of lane j as two f64 at unit q L + j, q < ceil(emax / 2)
//                   (rows shorter than that padded with +0.0)
//   octets          the LDS byte offsets (8 x slot) of the z values of entries 8r .. 8r + 7 of lane
//                   j as eight u16 at unit r L + j, r < ceil(emax / 8) (padding: 0, slot SW_ZERO)
// A sweep writes its results in its own order of the rows (steps, then waves, then lanes), 64
// consecutive doubles per wave; the caller moves them where they go with one parallel scatter.
// The first blocks of the teams are listed in a table of SW_TEAMS x SW_TWAVES records; a block past
// the last step is the empty block at the stream's end, whose zeros also serve every lane's pairs
// and octets past a block's bound.  A padded entry multiplies +0.0 by slot SW_ZERO's +0.0, and
// acc - (+0.0) is acc bit for bit, so the registered entries past a row's end need no select.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace femd_amd {

constexpr int SW_TEAMS = 4;                    // teams of waves, one step each in turn
constexpr int SW_TWAVES = 3;                   // waves per team
constexpr int SW_WAVES = SW_TEAMS * SW_TWAVES;
constexpr int SW_THREADS = 64 * SW_WAVES;
constexpr int SW_ROWS = 64 * SW_TWAVES;        // rows per step at most
constexpr int SW_SLOTS = 8192;                 // LDS z slots at most (64 KiB: byte offsets fit 16 bits)
constexpr int SW_ZERO = 0, SW_TRASH = 1;       // slot 0 holds 0.0; slot 1 takes values nobody reads
constexpr int SW_EREG = 24;                    // entries per row a wave holds in registers, at most
                                               // (the plan picks 8, 16 or 24 for its sweep)
constexpr int SW_TAIL = 128;                   // units of zeros past the last block

enum { SW_ILU_FWD = 0, SW_ILU_BWD = 1, SW_SSOR_FWD = 2, SW_SSOR_BWD = 3 };

constexpr bool sweep_has_diag(int kind) { return kind != SW_ILU_FWD; }
constexpr int sw_npairs(int emax) { return (emax + 1) / 2; }
constexpr int sw_noct(int emax) { return (emax + 7) / 8; }
// unit offsets inside a block of L lanes
constexpr int SW_HEAD = 1;                     // units of a block's header (next)
constexpr int64_t sw_u_pairs(int kind, int L) { return SW_HEAD + int64_t(L) + (sweep_has_diag(kind) ? int64_t(L) / 2 : 0); }
constexpr int64_t sw_u_octets(int kind, int L, int emax) { return sw_u_pairs(kind, L) + int64_t(sw_npairs(emax)) * L; }
constexpr int64_t sw_block_units(int kind, int L, int emax) { return sw_u_octets(kind, L, emax) + int64_t(sw_noct(emax)) * L; }
// byte offsets of lane j's row fields from the block's start
constexpr int64_t sw_b_vin(int j) { return 16 * (SW_HEAD + int64_t(j)); }
constexpr int64_t sw_b_sc(int j) { return sw_b_vin(j) + 12; }
constexpr int64_t sw_b_diag(int L, int j) { return 16 * (SW_HEAD + int64_t(L)) + 8 * int64_t(j); }

struct SweepPlan {
    int kind = 0;
    int n = 0;
    int nsteps = 0;
    int nlevels = 0;
    int nslots = 0;                  // slots used (incl. SW_ZERO, SW_TRASH)
    int max_emax = 0;
    std::vector<uint8_t> stream;     // the blocks (values, vin: zero) and SW_TAIL units of zeros
    std::vector<uint32_t> first;     // SW_TEAMS x SW_TWAVES records (u32 x 4 each): the teams' first blocks
    uint32_t zunit = 0;              // the empty block's first unit of zeros
    int ereg = SW_EREG;              // entries per row in registers (8, 16 or 24)
    std::vector<int64_t> vmap;       // per stored entry p of the pattern: its value's index (doubles) in
                                     // the stream, or -1 (its diagonal too, for kinds with diag)
    std::vector<int64_t> in_pos;     // per row: index (doubles) of its vin field
    std::vector<int> order;          // the rows in the sweep's order (where its results are written)
    int64_t entries = 0;             // stored entries the sweep reads (without padding)
    int64_t padded = 0;              // entry places in the pairs (with padding)
};

// Plan the sweep of `kind` over the square pattern (ip, ix) with diagonal positions diag: rows in
// ascending order (forward, the strictly lower entries) or descending (backward, the strictly
// upper ones).  Returns false, with the reason in why, when the schedule does not fit (more live
// values than slots, a stream beyond 16-byte offsets of 32 bits).
bool plan_sweep(int kind, int n, const int *ip, const int *ix, const int *diag, int max_rows, SweepPlan &out,
                std::string &why);

// The plan run on the host exactly as the GPU runs it (teams, blocks, slots), for tests: stream
// holds the packed values (vmap applied) and the vin fields; out (n values) receives the results in
// the sweep's order.  Returns false if a block's record does not name its wave's next block, or a
// read leaves the stream.
bool emulate_sweep(const SweepPlan &P, std::vector<uint8_t> &stream, double omega, double *out, std::string &why);

} // namespace femd_amd
