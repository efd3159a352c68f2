// grns_internal.hpp -- library-internal declarations shared by the host translation units.
#pragma once

#include <mutex>
#include <vector>

#include "../../include/grns.h"
#include "grns_kernels.hpp"

struct grns_ctx {
    int device = 0;
    int n = 0;
    int log2M = 0;
    grns::DevCtx h;                // host copy of the device constants
    grns::DevCtx *d = nullptr;     // device constants
    int *d_pow2 = nullptr;
    int *d_mrc = nullptr;
    double *d_mrc_d = nullptr;
    int *d_status = nullptr;
    unsigned *d_crt = nullptr;      // DevCtx::crt_words
    int max_blocks = 2048;         // arg-max stage-1 grid, from the occupancy query (max_grid_for)
    int max_blocks_occ = 2048;     // the same as set at context creation (test hooks restore it)
    // rare-item queues of the per-call eval / cmp kernels, one per stream (made on first use)
    struct RareQueue {
        void *stream;
        long long *items;
        unsigned *ctr;
        int *res;  // staged residues of the first queued items
    };
    std::mutex qmu;
    std::vector<RareQueue> queues;
    // optional timing of the first (streaming) kernel of eval / cmp / max: HIP events recorded on
    // the caller's stream around that kernel while enabled; pairs are reused round-robin and folded
    // into the total before reuse
    struct StageTimer {
        std::vector<void *> ev;    // 2 * slots events (start, end)
        std::vector<char> pending;  // slot holds an unread pair
        int next = 0;
        double ms = 0.0;
        long long launches = 0;
    };
    bool timing = false;
    bool reference_loops = false;  // division and scaling: the reference loops for every number (test hook)
    bool rare_staging = true;      // rare queues stage the residues of queued numbers (test hook)
    bool rare_spread = true;       // rare items spread over the sub-queues (test hook)
    int cus = 0;                   // compute units of the device
    int grid_cap = 0;              // streaming kernels: at most grid_cap x cus blocks (0: one per BLOCK numbers)
    StageTimer timers[3];
    std::vector<int> pow2;
    std::vector<int> mrc;
    std::vector<double> mrc_d;
    std::vector<unsigned> crt;
};

namespace grns {
inline int groups_for(int n) {
    int G = 1;
    while (G < n) G <<= 1;
    return G;
}
}  // namespace grns

// Lanes per number of the word division: a power of two with room for every residue and every
// word of the CRT sum (0: no group of 64 lanes has room).
inline int words_group(const grns_ctx *c) {
    int G = grns::groups_for(c->n);
    while (G < c->h.words) G <<= 1;
    return G <= 64 ? G : 0;
}

namespace grns {
}  // namespace grns
