// Direct solve on the coarsest multigrid level (kernels_coarse.hip).
#pragma once

#include <cstdint>

#include "tng_internal.hpp"

namespace tng {

struct CoarseDirect {
    static constexpr int kMaxUnknowns = 6000;  // 288 MB of FP64 inverse at most
    int nu = 0, n = 0;                          // unique nodes, unknowns (3 nu)
    int64_t ncopies = 0;
    bool ready = false;
    DevBuf d_uid, d_rep, d_M, d_row, d_col, d_b, d_x;
    void numbering(const tng_domain &m);
    // the inverse of A over the unique node components; x, y: FP64 work vectors on A's mesh
    // (fp32_levels: the V-cycle's other levels are FP32; with the fs/fs penalty
    // the rotations are then projected out, see project_rotations)
    void build(tng_operator &A, tng_vector &x, tng_vector &y, bool fp32_levels, hipStream_t s);
    // e0 = A^{-1} r0 on every node copy (r0 consistent over the copies)
    void solve(tng_vector &e0, const tng_vector &r0, hipStream_t s);
    // M <- P M P, P the projection off the three rigid rotations (fs/fs penalty)
    void project_rotations(const tng_domain &m, hipStream_t s);
};

}  // namespace tng
