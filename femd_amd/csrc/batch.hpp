// batch.hpp -- the batch handle: independent square systems concatenated into one block-diagonal
// matrix whose pattern has one segment per system, solved by the inner CG engine (batch.hip).
// Also the groups of the serving loop (serve.hip).
#pragma once
#include "cg_engine.hpp"

struct femd_batch_s {
    int np = 0;
    int64_t nrows = 0, nnz = 0, max_rows = 0;
    std::vector<femd_csr> members;
    std::vector<int64_t> roff, voff;            // row and value offsets (np + 1)
    std::unique_ptr<femd_csr_s> A;              // the block-diagonal matrix, one segment per system
    femd_amd::CGEngine engine[2];               // FP32 and FP64 inner iterations
    bool reorder = true;                        // the engines work in the pattern's LocalOrder
};

namespace femd_amd {
// Copy the members' patterns and values into a new batch (work on stream s, synchronized); its
// engines work in the inner CG's own order of the unknowns when `reorder`.
std::unique_ptr<femd_batch_s> batch_build(int np, const femd_csr *mats, hipStream_t s, bool reorder = true);
// Re-copy the members' values (after a re-assembly).
void batch_copy_values(femd_batch_s *B, hipStream_t s);
} // namespace femd_amd
