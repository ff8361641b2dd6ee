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
// step that reads it (on C1's RCM order at most ~5.7k values are live at once); the LDS the slots
// leave is the ring.  Everything a step
// reads besides those slots -- its header, the rows' records and their entries (value, slot of the
// column) -- is one contiguous piece of a static byte stream in HBM, which producer waves copy into
// an LDS ring ahead of use while the consumer waves compute.  A consumer wave reads the next
// step's piece into registers during the current step, so that after a barrier it only gathers the
// z values its rows need and sums:
//
//   step:  header (SW_HDR bytes): u32 size (bytes of the step), u32 waves, u32 wb[s + 4],
//          u32 wb[s + 5] (the blocks the producer of step s loads for step s + 4), then per wave w a
//          record of four u32 (woff, emax, L, npre); then one block per consumer wave with rows, at
//          the step's offset + woff, for its L lanes (its rows, longest first, rounded up to a
//          multiple of 8), every section a multiple of 16 bytes:
//            f64 vin[L]    the row's input: r (forward, filled per application), y (backward,
//                          written by the forward sweep)
//            i32 dest[L]   where the result goes: forward -- the backward stream's vin field of the
//                          row (index in doubles), backward -- the output vector's index; -1: none
//            u32 sc[L]     the row's own LDS slot (SW_TRASH when no later row reads it) | its
//                          entry count << 16
//            f64 diag[L]   the row's diagonal (kinds that divide by it)
//            u32 pre[npre] padded to a multiple of 4: where (in 16-byte units from the block's start)
//                          pair q of the entries starts (q < npairs = ceil(emax / 2)), then where
//                          slot octet r starts (r < ceil(emax / 8))
//            pairs         entries 2q and 2q + 1 of lane j as two f64 at pre[q] + j (only the
//                          rows with more than 2q entries: a jagged-diagonal layout by pairs)
//            octets        the LDS slots of entries 8r .. 8r + 7 of lane j as eight u16 at
//                          pre[npairs + r] + j (rows with more than 8r entries)
//
// Producer wave p handles the steps s with s % 4 == p: at step s it stores the blocks it loaded
// four steps earlier -- blocks [wb[s], wb[s+1]) of SW_BLOCK bytes -- into the ring and loads those
// of step s + 4, whose range step s's header holds (the first four ranges come from wb in HBM).
// The plan places wb so that, while step s runs, the ring holds step s's bytes and all of step
// s + 1's, and no store overwrites bytes the step may still read.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace femd_amd {

constexpr int SW_CWAVES = 8;                   // consumer waves
constexpr int SW_PWAVES = 4;                   // producer waves
constexpr int SW_THREADS = 64 * (SW_CWAVES + SW_PWAVES);
constexpr int SW_ROWS = 64 * SW_CWAVES;        // rows per step at most
constexpr int SW_BLOCK = 1024;                 // bytes per producer block (64 lanes x 16 B)
constexpr int SW_LDS = 152 * 1024;             // LDS bytes: the z slots, then the ring
constexpr int SW_SLOTS = 8192;                 // LDS z slots at most
constexpr int SW_ZERO = 0, SW_TRASH = 1;       // slot 0 holds 0.0; slot 1 takes values nobody reads
constexpr int SW_PMAX = 20;                    // blocks a producer wave carries per turn
constexpr int SW_HDR = 144;
constexpr int SW_AHEAD = SW_PWAVES;            // steps between a block's load and its store

enum { SW_ILU_FWD = 0, SW_ILU_BWD = 1, SW_SSOR_FWD = 2, SW_SSOR_BWD = 3 };

constexpr int SW_EREG = 24;                    // entries per row a consumer keeps in registers
constexpr int SW_OREG = SW_EREG / 8;           // slot octets in registers

constexpr bool sweep_has_diag(int kind) { return kind != SW_ILU_FWD; }
constexpr int sw_npairs(int emax) { return (emax + 1) / 2; }
constexpr int sw_noct(int emax) { return (emax + 7) / 8; }
constexpr int sw_npre(int emax) { return sw_npairs(emax) + sw_noct(emax); }
// byte offsets inside a wave block of L lanes
constexpr int64_t sw_off_dest(int L) { return 8 * int64_t(L); }
constexpr int64_t sw_off_sc(int L) { return 12 * int64_t(L); }
constexpr int64_t sw_off_diag(int L) { return 16 * int64_t(L); }
constexpr int64_t sw_off_pre(int kind, int L) { return (sweep_has_diag(kind) ? 24 : 16) * int64_t(L); }
// the first pair (16-byte units from the block's start)
constexpr int64_t sw_unit_pairs(int kind, int L, int emax) { return (sw_off_pre(kind, L) + 4 * ((sw_npre(emax) + 3) / 4 * 4)) / 16; }
// a wave block with T2 pair units and T8 octet units
constexpr int64_t sw_wave_bytes(int kind, int L, int emax, int64_t T2, int64_t T8)
{
    return 16 * (sw_unit_pairs(kind, L, emax) + T2 + T8);
}

struct SweepPlan {
    int kind = 0;
    int n = 0;
    int nsteps = 0;                  // steps the kernel runs (idle ones included)
    int nreal = 0;                   // steps with rows
    int nlevels = 0;
    int nslots = 0;                  // slots used (incl. SW_ZERO, SW_TRASH)
    int max_emax = 0;
    int zbytes = 0;                  // LDS bytes of the z slots (whole blocks)
    int rbytes = 0;                  // LDS bytes of the ring (whole blocks): the rest
    std::vector<uint8_t> stream;     // the static bytes (values, vin: zero), padded to whole blocks,
                                     // and one spare block
    std::vector<int> wb;             // nsteps + 2 * SW_AHEAD + 2 entries: blocks in the ring before step s
    std::vector<int64_t> vmap;       // per stored entry p of the pattern: its value's index (doubles) in
                                     // the stream, or -1 (its diagonal too, for kinds with diag)
    std::vector<int64_t> in_pos;     // per row: index (doubles) of its vin field
    int64_t entries = 0;             // stored entries the sweep reads (without padding)
    int64_t nblocks() const { return static_cast<int64_t>(stream.size()) / SW_BLOCK; }
};

// Plan the sweep of `kind` over the square pattern (ip, ix) with diagonal positions diag: rows in
// ascending order (forward, the strictly lower entries) or descending (backward, the strictly
// upper ones).  dest[i]: the dest field of row i (see above).  Returns false, with the reason in
// why, when the schedule does not fit the LDS (more live values than slots, or a step larger than
// the ring allows).
bool plan_sweep(int kind, int n, const int *ip, const int *ix, const int *diag, const std::vector<int> &dest,
                int max_rows, SweepPlan &out, std::string &why);

// The plan run on the host exactly as the GPU runs it (ring, slots, producer schedule), for tests:
// stream holds the packed values (vmap applied) and the vin fields; out receives the dest writes.
// Returns false if the run reads a ring byte that no producer stored first (a planning error).
bool emulate_sweep(const SweepPlan &P, std::vector<uint8_t> &stream, double omega, double *out, int64_t nout,
                   std::string &why);

} // namespace femd_amd
