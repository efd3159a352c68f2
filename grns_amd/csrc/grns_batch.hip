// grns_batch.hip -- batched form: many arrays of uneven length (and several operations of one
// moduli set) in one persistent launch.  Work is cut into tiles of `tile` numbers; persistent
// blocks take tiles from an atomic counter (work stealing across tasks, so short and long arrays
// mix freely), look up the task of the tile by binary search over the tiles' prefix, and run the
// same per-number code as the per-call kernels, so every result is identical to theirs.  The
// counter is reset by a one-thread kernel in the same stream, so the whole sequence can be
// captured in a hipGraph and replayed.  An autotuner times candidate tile sizes, persistent
// blocks per CU and the lane mapping (one lane or one group of lanes per number) on the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/grns_debug.h"
#include "grns_internal.hpp"

using namespace grns;

namespace grns {

struct BatchTask {
    int op;
    int pad;
    long long count;
    long long tile0;  // first tile index of this task within its list
    const int *a;
    long long asm_, asn;
    const int *b;
    long long bsm, bsn;
    void *o0;
    void *o1;
    const unsigned *D;
    MpintArgs mp;
};

// The task holding tile t: the last task whose first tile is <= t (tasks[0].tile0 = 0 <= t and
// the first tiles increase strictly: every task has a tile).  Wave 0 searches 64-ary: each lane
// tests one candidate and a ballot keeps the last one at or below t, so 4096 tasks take two
// rounds of dependent loads instead of twelve.
GRNS_D int find_task_wave(const BatchTask *__restrict__ tasks, int ntasks, long long t) {
    const int lane = threadIdx.x & 63;
    int lo = 0, span = ntasks;
    while (span > 1) {
        const int step = (span + 63) / 64;
        const int idx = lo + lane * step;
        const bool le = lane * step < span && tasks[idx].tile0 <= t;
        const unsigned long long m = __ballot(le);  // bit 0 set: tasks[lo].tile0 <= t
        const int last = 63 - __clzll(m);
        const int nlo = lo + last * step;
        span = (lo + span - nlo) < step ? (lo + span - nlo) : step;
        lo = nlo;
    }
    return lo;
}

// Fetch the next tile for the block: returns false when the work is exhausted.
GRNS_D bool next_tile(const BatchTask *__restrict__ tasks, int ntasks, long long total, unsigned long long *counter,
                      long long &t, int &k) {
    __shared__ long long s_t;
    __shared__ int s_k;
    if (threadIdx.x < 64) {
        unsigned long long v = 0;
        if (threadIdx.x == 0) v = atomicAdd(counter, 1ull);
        v = __shfl(v, 0, 64);
        const int kk = (long long)v < total ? find_task_wave(tasks, ntasks, (long long)v) : 0;
        if (threadIdx.x == 0) {
            s_t = (long long)v;
            s_k = kk;
        }
    }
    __syncthreads();
    // the same for the whole block: keep them in scalar registers (task fields become scalar loads)
    const unsigned long long tv = (unsigned long long)s_t;
    t = (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(tv >> 32)) << 32) |
                    (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)tv));
    k = __builtin_amdgcn_readfirstlane(s_k);
    __syncthreads();
    return t < total;
}

// Operation class of a thread-mapped task list: one kernel instantiation per class, so that each
// carries only its class's code and runs with its registers and occupancy (as the per-call kernels
// do): evaluation (and its fast form), comparison, and the mpint add / sub / mul, whose codes need
// about the same registers (104-114 VGPRs at n = 32) and so share one kernel and one tile space.
enum : int { OPS_EVAL = 1, OPS_CMP = 2, OPS_MPINT = 4 };
constexpr int OP_CLASSES = 3;
constexpr int BATCH_COUNTERS = 4;  // tile counters: one per thread list (3), one for the group list

template <int N, int OPS>
__global__ __launch_bounds__(BLOCK) void k_batch_thread(const DevCtx *__restrict__ c, const BatchTask *__restrict__ tasks,
                                                        int ntasks, long long total, int tile,
                                                        unsigned long long *counter, int *__restrict__ status) {
    long long t;
    int k;
    while (next_tile(tasks, ntasks, total, counter, t, k)) {
        const BatchTask &T = tasks[k];
        long long e0 = (t - T.tile0) * tile;
        long long e1 = e0 + tile < T.count ? e0 + tile : T.count;
        switch (T.op) {
            case GRNS_OP_EVAL:
            case GRNS_OP_EVAL_FAST:
                if constexpr ((OPS & OPS_EVAL) != 0) {
                    interval_t *o = reinterpret_cast<interval_t *>(T.o0);
                    for (long long base = e0; base < e1; base += BLOCK) {
                        if (T.op == GRNS_OP_EVAL)
                            eval_step<N, false, LD_MM>(c, T.a, T.asm_, 1, base + threadIdx.x, e1, o, status);
                        else eval_step<N, true, LD_MM>(c, T.a, T.asm_, 1, base + threadIdx.x, e1, o, status);
                    }
                }
                break;
            case GRNS_OP_CMP:
                if constexpr ((OPS & OPS_CMP) != 0) {
                    for (long long base = e0; base < e1; base += BLOCK)
                        cmp_step<N, LD_MM>(c, T.a, T.asm_, 1, T.b, T.bsm, 1, base + threadIdx.x, e1,
                                           reinterpret_cast<int *>(T.o0), status);
                }
                break;
            case GRNS_OP_MPINT_ADD:
                if constexpr ((OPS & OPS_MPINT) != 0)
                    for (long long j = e0 + threadIdx.x; j < e1; j += BLOCK) mpint_step<N, 0, true>(c, T.mp, j);
                break;
            case GRNS_OP_MPINT_SUB:
                if constexpr ((OPS & OPS_MPINT) != 0)
                    for (long long j = e0 + threadIdx.x; j < e1; j += BLOCK) mpint_step<N, 1, true>(c, T.mp, j);
                break;
            case GRNS_OP_MPINT_MUL:
                if constexpr ((OPS & OPS_MPINT) != 0)
                    for (long long j = e0 + threadIdx.x; j < e1; j += BLOCK) mpint_step<N, 2, true>(c, T.mp, j);
                break;
            default:
                break;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_batch_group(const DevCtx *__restrict__ c, int G,
                                                       const BatchTask *__restrict__ tasks, int ntasks, long long total,
                                                       int tile, unsigned long long *counter, int *__restrict__ status) {
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    const long long wave = threadIdx.x >> 6;
    const long long gid = (threadIdx.x & 63) / G;
    long long t;
    int k;
    while (next_tile(tasks, ntasks, total, counter, t, k)) {
        const BatchTask &T = tasks[k];
        long long e0 = (t - T.tile0) * tile;
        long long e1 = e0 + tile < T.count ? e0 + tile : T.count;
        for (long long j0 = e0 + wave * P; j0 < e1; j0 += (BLOCK / 64) * P) {
            long long j = j0 + gid;
            if (j >= e1) continue;
            int st = 0;
            const bool act = g.gl < g.L;
            long long oa = j * T.asn + (long long)g.gl * T.asm_;
            switch (T.op) {
                case GRNS_OP_EVAL:
                case GRNS_OP_EVAL_FAST: {
                    int xi = act ? T.a[oa] : 0;
                    interval_t r = (T.op == GRNS_OP_EVAL) ? group_eval<false>(c, g, xi, st) : group_eval<true>(c, g, xi, st);
                    if (g.gl == 0) store_interval(reinterpret_cast<interval_t *>(T.o0) + j, r);
                    break;
                }
                case GRNS_OP_CMP: {
                    int xi = act ? T.a[oa] : 0;
                    int yi = act ? T.b[j * T.bsn + (long long)g.gl * T.bsm] : 0;
                    int r = group_cmp(c, g, xi, yi, st);
                    if (g.gl == 0) reinterpret_cast<int *>(T.o0)[j] = r;
                    break;
                }
                case GRNS_OP_MPINT_ADD: group_mpint<0>(c, g, T.mp, j); break;
                case GRNS_OP_MPINT_SUB: group_mpint<1>(c, g, T.mp, j); break;
                case GRNS_OP_MPINT_MUL: group_mpint<2>(c, g, T.mp, j); break;
                case GRNS_OP_DIV:
                case GRNS_OP_DIV_FAST: {
                    int xi = act ? T.a[oa] : 0;
                    int di = act ? T.b[oa] : 0;
                    int q, r;
                    if (g.G >= c->words && g.G >= 4 && div_words_ok(c, g, xi, di)) {
                        switch (g.G) {
                            case 4: group_div_words<4>(c, g, xi, di, q, r); break;
                            case 8: group_div_words<8>(c, g, xi, di, q, r); break;
                            case 16: group_div_words<16>(c, g, xi, di, q, r); break;
                            case 32: group_div_words<32>(c, g, xi, di, q, r); break;
                            default: group_div_words<64>(c, g, xi, di, q, r); break;
                        }
                    } else if (T.op == GRNS_OP_DIV) {
                        group_div<false>(c, g, xi, di, q, r, st);
                    } else {
                        group_div<true>(c, g, xi, di, q, r, st);
                    }
                    if (act) {
                        reinterpret_cast<int *>(T.o0)[oa] = q;
                        reinterpret_cast<int *>(T.o1)[oa] = r;
                    }
                    break;
                }
                case GRNS_OP_SCALE2POW: {
                    int xi = act ? T.a[oa] : 0;
                    const unsigned Dj = T.D[j];
                    bool ok = false;
                    int v = 0;
                    if (g.G >= c->words && g.G >= 4 && group_all(g, !act || (unsigned)xi < (unsigned)c->m[g.gl])) {
                        switch (g.G) {
                            case 4: v = group_scale2pow_words<4>(c, g, xi, Dj, ok); break;
                            case 8: v = group_scale2pow_words<8>(c, g, xi, Dj, ok); break;
                            case 16: v = group_scale2pow_words<16>(c, g, xi, Dj, ok); break;
                            case 32: v = group_scale2pow_words<32>(c, g, xi, Dj, ok); break;
                            default: v = group_scale2pow_words<64>(c, g, xi, Dj, ok); break;
                        }
                    }
                    if (!ok) v = group_scale2pow(c, g, xi, Dj);
                    if (act) reinterpret_cast<int *>(T.o0)[oa] = v;
                    break;
                }
                default:
                    break;
            }
            if (g.gl == 0 && st) atomicOr(status, st);
        }
    }
}

__global__ void k_batch_reset(unsigned long long *counters) {
    if (threadIdx.x < BATCH_COUNTERS) counters[threadIdx.x] = 0;
}

}  // namespace grns

namespace {

bool thread_kernel_available(int n) { return n == 4 || n == 8 || n == 16 || n == 32 || n == 64; }

bool elementwise(int op) { return op >= GRNS_OP_EVAL && op <= GRNS_OP_MPINT_MUL; }

// Thread lists address digit rows by 32-bit lane offsets from scalar row pointers: numbers stored
// digit-major (number stride 1) and at most 2^29 of them.  Other layouts go to the group list.
bool lane_layout(const grns_batch_task &u) {
    if (u.count > ((size_t)1 << 29)) return false;
    if (u.op == GRNS_OP_EVAL || u.op == GRNS_OP_EVAL_FAST) return u.a_num_stride == 1;
    if (u.op == GRNS_OP_CMP) return u.a_num_stride == 1 && u.b_num_stride == 1;
    return true;
}

BatchTask to_dev(const grns_batch_task &u) {
    BatchTask t;
    std::memset(&t, 0, sizeof(t));
    t.op = u.op;
    t.count = (long long)u.count;
    t.a = u.a; t.asm_ = u.a_mod_stride; t.asn = u.a_num_stride;
    t.b = u.b; t.bsm = u.b_mod_stride; t.bsn = u.b_num_stride;
    t.o0 = u.out0; t.o1 = u.out1; t.D = u.D;
    t.mp.xd = u.mx.digits; t.mp.xs = u.mx.sign; t.mp.xe = reinterpret_cast<const er_t *>(u.mx.eval);
    t.mp.yd = u.my.digits; t.mp.ys = u.my.sign; t.mp.re = reinterpret_cast<er_t *>(u.mz.eval);
    t.mp.ye = reinterpret_cast<const er_t *>(u.my.eval);
    t.mp.rd = u.mz.digits; t.mp.rs = u.mz.sign;
    t.mp.ovf = u.overflow; t.mp.ld = u.mx.ld;
    return t;
}

// Tasks split by mapping and operation: the thread-per-number lists first, one per operation, then
// the group-per-number list; each list has its own tile prefix (tile0 counts from 0 within it).
struct ThreadList {
    int ops;          // OPS_* of the list's operation
    int first;        // index of its first task in TaskPlan::all
    int count;        // tasks
    long long tiles;  // tiles of `tile` numbers
};

struct TaskPlan {
    std::vector<BatchTask> all;
    std::vector<ThreadList> lists;
    int nthr = 0, ngrp = 0;
    long long grp_tiles = 0;
    long long grp_tile = 1;  // numbers per tile of the group list: a group does far more work per number
};

int op_class(int op) {
    switch (op) {
        case GRNS_OP_EVAL:
        case GRNS_OP_EVAL_FAST: return OPS_EVAL;
        case GRNS_OP_CMP: return OPS_CMP;
        default: return OPS_MPINT;
    }
}

void make_plan(const grns_batch_task *u, int n, int ctx_n, const grns_batch_config &cfg, TaskPlan &P) {
    P.all.clear();
    P.lists.clear();
    std::vector<BatchTask> grp, by_op[OP_CLASSES];
    const bool thread_ok = thread_kernel_available(ctx_n) && cfg.threads_per_number == 1;
    const long long tile = cfg.tile;
    P.grp_tile = tile / 16 > 0 ? tile / 16 : 1;
    P.grp_tiles = 0;
    for (int i = 0; i < n; i++) {
        if (u[i].count == 0 || u[i].op == GRNS_OP_MAX) continue;
        BatchTask t = to_dev(u[i]);
        if (thread_ok && elementwise(u[i].op) && lane_layout(u[i])) {
            by_op[__builtin_ctz(op_class(u[i].op))].push_back(t);
        } else {
            t.tile0 = P.grp_tiles;
            P.grp_tiles += (t.count + P.grp_tile - 1) / P.grp_tile;
            grp.push_back(t);
        }
    }
    for (int k = 0; k < OP_CLASSES; k++) {
        if (by_op[k].empty()) continue;
        ThreadList L{1 << k, (int)P.all.size(), (int)by_op[k].size(), 0};
        for (BatchTask &t : by_op[k]) {
            t.tile0 = L.tiles;
            L.tiles += (t.count + tile - 1) / tile;
            P.all.push_back(t);
        }
        P.lists.push_back(L);
    }
    P.nthr = (int)P.all.size();
    P.ngrp = (int)grp.size();
    P.all.insert(P.all.end(), grp.begin(), grp.end());
}

// Counter reset plus one persistent launch per non-empty list, in order on the stream; d_tasks
// must hold P.all.
template <int NN>
void launch_thread_list(grns_ctx *c, const ThreadList &L, const BatchTask *d_tasks, unsigned long long *counter,
                        int tile, dim3 grid, hipStream_t s) {
    const dim3 blk(BLOCK);
    const BatchTask *tl = d_tasks + L.first;
    switch (L.ops) {
        case OPS_EVAL:
            hipLaunchKernelGGL((k_batch_thread<NN, OPS_EVAL>), grid, blk, 0, s, c->d, tl, L.count, L.tiles, tile, counter,
                               c->d_status);
            break;
        case OPS_CMP:
            hipLaunchKernelGGL((k_batch_thread<NN, OPS_CMP>), grid, blk, 0, s, c->d, tl, L.count, L.tiles, tile, counter,
                               c->d_status);
            break;
        default:
            hipLaunchKernelGGL((k_batch_thread<NN, OPS_MPINT>), grid, blk, 0, s, c->d, tl, L.count, L.tiles, tile,
                               counter, c->d_status);
            break;
    }
}

grns_status enqueue_plan(grns_ctx *c, const TaskPlan &P, const BatchTask *d_tasks, unsigned long long *d_counters,
                         const grns_batch_config &cfg, int cus, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, d_counters);
    dim3 grid((unsigned)(cus * cfg.blocks_per_cu)), blk(BLOCK);
    for (size_t li = 0; li < P.lists.size(); li++) {
        unsigned long long *counter = d_counters + li;
        switch (c->n) {
            case 4: launch_thread_list<4>(c, P.lists[li], d_tasks, counter, cfg.tile, grid, s); break;
            case 8: launch_thread_list<8>(c, P.lists[li], d_tasks, counter, cfg.tile, grid, s); break;
            case 16: launch_thread_list<16>(c, P.lists[li], d_tasks, counter, cfg.tile, grid, s); break;
            case 32: launch_thread_list<32>(c, P.lists[li], d_tasks, counter, cfg.tile, grid, s); break;
            case 64: launch_thread_list<64>(c, P.lists[li], d_tasks, counter, cfg.tile, grid, s); break;
            default: return GRNS_ERR_UNSUPPORTED;
        }
    }
    if (P.ngrp) {
        // group width: the word division's when it has one (the group tasks are mostly divisions)
        const int G = words_group(c) ? words_group(c) : groups_for(c->n);
        hipLaunchKernelGGL(k_batch_group, grid, blk, 0, s, c->d, G, d_tasks + P.nthr, P.ngrp,
                           P.grp_tiles, (int)P.grp_tile, d_counters + BATCH_COUNTERS - 1, c->d_status);
    }
    return hipGetLastError() == hipSuccess ? GRNS_OK : GRNS_ERR_HIP;
}

}  // namespace

struct grns_batch {
    grns_ctx *ctx = nullptr;
    std::vector<grns_batch_task> user;
    TaskPlan plan;
    BatchTask *d_tasks = nullptr;
    unsigned long long *d_counters = nullptr;
    int cap = 0;
    grns_batch_config cfg{1024, 4, 1, 0};
    int cus = 256;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(grns_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

grns_status rebuild(grns_batch *b) {
    drop_graph(b);
    make_plan(b->user.data(), (int)b->user.size(), b->ctx->n, b->cfg, b->plan);
    if (!b->plan.all.empty() && hipMemcpy(b->d_tasks, b->plan.all.data(), sizeof(BatchTask) * b->plan.all.size(),
                                          hipMemcpyHostToDevice) != hipSuccess)
        return GRNS_ERR_HIP;
    return GRNS_OK;
}

grns_status enqueue(grns_batch *b, hipStream_t s) {
    return enqueue_plan(b->ctx, b->plan, b->d_tasks, b->d_counters, b->cfg, b->cus, s);
}

// Every pointer a task's operation reads or writes is set, and mpint arrays are as long as the
// task (the checks of the per-call entry points).  The serving form (allow_max) also takes the
// arg-max.  A task of count 0 is valid whatever its pointers.
bool task_valid(const grns_batch_task &u, bool allow_max) {
    if (u.op < GRNS_OP_EVAL || u.op > (allow_max ? GRNS_OP_MAX : GRNS_OP_SCALE2POW)) return false;
    if (u.count == 0) return true;
    switch (u.op) {
        case GRNS_OP_MPINT_ADD:
        case GRNS_OP_MPINT_SUB:
        case GRNS_OP_MPINT_MUL: {
            const grns_mpint_array *m[3] = {&u.mx, &u.my, &u.mz};
            for (const grns_mpint_array *a : m)
                if (!a->digits || !a->sign || !a->eval || a->ld != u.mx.ld) return false;
            return u.mx.ld >= (long long)u.count;
        }
        case GRNS_OP_CMP: return u.a && u.b && u.out0;
        case GRNS_OP_DIV:
        case GRNS_OP_DIV_FAST: return u.a && u.b && u.out0 && u.out1;
        case GRNS_OP_SCALE2POW: return u.a && u.D && u.out0;
        case GRNS_OP_MAX: return u.a && u.out0 && u.out1;
        default: return u.a && u.out0;  // eval, eval_fast
    }
}

}  // namespace

extern "C" {

grns_status grns_batch_create(grns_ctx *ctx, const grns_batch_task *tasks, int ntasks, grns_batch **out) {
    if (!ctx || !out || ntasks < 0 || (ntasks && !tasks) || ctx->device < 0 || ctx->n > 64) return GRNS_ERR_ARG;
    for (int i = 0; i < ntasks; i++)
        if (!task_valid(tasks[i], false)) return GRNS_ERR_ARG;  // arg-max: use grns_max
    grns_batch *b = new (std::nothrow) grns_batch();
    if (!b) return GRNS_ERR_OOM;
    b->ctx = ctx;
    b->user.assign(tasks, tasks + ntasks);
    b->cap = ntasks > 0 ? ntasks : 1;
    int dev = ctx->device;
    (void)hipSetDevice(dev);
    if (hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) b->cus = 256;
    if (!thread_kernel_available(ctx->n)) b->cfg.threads_per_number = 0;
    if (hipMalloc(&b->d_tasks, sizeof(BatchTask) * b->cap) != hipSuccess ||
        hipMalloc(&b->d_counters, BATCH_COUNTERS * sizeof(unsigned long long)) != hipSuccess) {
        grns_batch_destroy(b);
        return GRNS_ERR_OOM;
    }
    grns_status st = rebuild(b);
    if (st != GRNS_OK) {
        grns_batch_destroy(b);
        return st;
    }
    *out = b;
    return GRNS_OK;
}

grns_status grns_batch_set_config(grns_batch *b, const grns_batch_config *cfg) {
    if (!b || !cfg || cfg->tile < 1 || cfg->blocks_per_cu < 1 || cfg->blocks_per_cu > 64) return GRNS_ERR_ARG;
    grns_batch_config nc = *cfg;
    if (!thread_kernel_available(b->ctx->n)) nc.threads_per_number = 0;
    if (nc.threads_per_number != 0) nc.threads_per_number = 1;
    bool reshape = nc.tile != b->cfg.tile || nc.threads_per_number != b->cfg.threads_per_number;
    bool regraph = reshape || nc.blocks_per_cu != b->cfg.blocks_per_cu || nc.use_graph != b->cfg.use_graph;
    b->cfg = nc;
    if (reshape) return rebuild(b);
    if (regraph) drop_graph(b);
    return GRNS_OK;
}

grns_status grns_batch_get_config(const grns_batch *b, grns_batch_config *cfg) {
    if (!b || !cfg) return GRNS_ERR_ARG;
    *cfg = b->cfg;
    return GRNS_OK;
}

grns_status grns_debug_batch_per_call(grns_batch *b, void *stream) {
    if (!b) return GRNS_ERR_ARG;
    grns_ctx *c = b->ctx;
    for (const grns_batch_task &u : b->user) {
        if (u.count == 0) continue;
        grns_status st = GRNS_OK;
        switch (u.op) {
            case GRNS_OP_EVAL:
            case GRNS_OP_EVAL_FAST:
                st = grns_eval(c, u.a, u.a_mod_stride, u.a_num_stride, u.count, static_cast<grns_interval *>(u.out0),
                               u.op == GRNS_OP_EVAL_FAST, stream);
                break;
            case GRNS_OP_CMP:
                st = grns_cmp(c, u.a, u.a_mod_stride, u.a_num_stride, u.b, u.b_mod_stride, u.b_num_stride, u.count,
                              static_cast<int *>(u.out0), stream);
                break;
            case GRNS_OP_MPINT_ADD: st = grns_mpint_add(c, &u.mx, &u.my, &u.mz, u.count, u.overflow, stream); break;
            case GRNS_OP_MPINT_SUB: st = grns_mpint_sub(c, &u.mx, &u.my, &u.mz, u.count, u.overflow, stream); break;
            case GRNS_OP_MPINT_MUL: st = grns_mpint_mul(c, &u.mx, &u.my, &u.mz, u.count, u.overflow, stream); break;
            case GRNS_OP_DIV:
            case GRNS_OP_DIV_FAST:
                st = grns_div(c, u.a, u.b, u.a_mod_stride, u.a_num_stride, u.count, static_cast<int *>(u.out0),
                              static_cast<int *>(u.out1), u.op == GRNS_OP_DIV_FAST ? GRNS_DIV_FAST : 0, stream);
                break;
            case GRNS_OP_SCALE2POW:
                st = grns_scale2pow(c, u.a, u.a_mod_stride, u.a_num_stride, u.D, u.count, static_cast<int *>(u.out0),
                                    stream);
                break;
            default: st = GRNS_ERR_ARG; break;
        }
        if (st != GRNS_OK) return st;
    }
    return GRNS_OK;
}

grns_status grns_batch_launch(grns_batch *b, void *stream) {
    if (!b) return GRNS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    (void)hipSetDevice(b->ctx->device);
    if (b->plan.all.empty()) return GRNS_OK;
    if (!b->cfg.use_graph) return enqueue(b, s);
    if (!b->exec) {
        hipStream_t cs;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return GRNS_ERR_HIP;
        grns_status st = GRNS_OK;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) st = GRNS_ERR_HIP;
        if (st == GRNS_OK) st = enqueue(b, cs);
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(cs, &g) != hipSuccess) st = GRNS_ERR_HIP;
        (void)hipStreamDestroy(cs);
        if (st != GRNS_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        b->graph = g;
        if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            drop_graph(b);
            return GRNS_ERR_HIP;
        }
    }
    return hipGraphLaunch(b->exec, s) == hipSuccess ? GRNS_OK : GRNS_ERR_HIP;
}

grns_status grns_batch_tune(grns_batch *b, void *stream, int reps, grns_batch_config *best, double *best_ms) {
    if (!b || reps < 1) return GRNS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    grns_batch_config keep = b->cfg, top = b->cfg;
    double top_ms = -1;
    const int tiles[] = {256, 1024, 4096, 16384};
    const int bpcs[] = {1, 2, 4, 8};
    std::vector<int> maps;
    bool any_elem = false;
    for (const auto &u : b->user) any_elem |= elementwise(u.op);
    maps.push_back(0);
    if (any_elem && thread_kernel_available(b->ctx->n)) maps.push_back(1);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return GRNS_ERR_HIP;
    grns_status st = GRNS_OK;
    for (int m : maps) {
        for (int tl : tiles) {
            for (int bp : bpcs) {
                grns_batch_config c{tl, bp, m, keep.use_graph};
                if ((st = grns_batch_set_config(b, &c)) != GRNS_OK) break;
                if ((st = grns_batch_launch(b, s)) != GRNS_OK) break;  // warm-up (and graph build)
                (void)hipEventRecord(e0, s);
                for (int r = 0; r < reps && st == GRNS_OK; r++) st = grns_batch_launch(b, s);
                (void)hipEventRecord(e1, s);
                if (hipEventSynchronize(e1) != hipSuccess) st = GRNS_ERR_HIP;
                if (st != GRNS_OK) break;
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                double per = ms / reps;
                if (top_ms < 0 || per < top_ms) {
                    top_ms = per;
                    top = c;
                }
            }
            if (st != GRNS_OK) break;
        }
        if (st != GRNS_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != GRNS_OK) {
        grns_batch_set_config(b, &keep);
        return st;
    }
    st = grns_batch_set_config(b, &top);
    if (best) *best = top;
    if (best_ms) *best_ms = top_ms;
    return st;
}

grns_status grns_batch_destroy(grns_batch *b) {
    if (!b) return GRNS_ERR_ARG;
    drop_graph(b);
    if (b->d_tasks) (void)hipFree(b->d_tasks);
    if (b->d_counters) (void)hipFree(b->d_counters);
    delete b;
    return GRNS_OK;
}

}  // extern "C"

// ============================================================================ serving form
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

namespace {

using Clock = std::chrono::steady_clock;

size_t task_bytes(const grns_batch_task &t, int n) {
    const size_t c = t.count, nn = (size_t)n;
    switch (t.op) {
        case GRNS_OP_EVAL:
        case GRNS_OP_EVAL_FAST: return c * (4 * nn + 32);
        case GRNS_OP_CMP: return c * (8 * nn + 4);
        case GRNS_OP_MPINT_ADD:
        case GRNS_OP_MPINT_SUB:
        case GRNS_OP_MPINT_MUL: return c * 3 * (4 * nn + 36);
        case GRNS_OP_DIV:
        case GRNS_OP_DIV_FAST: return c * 16 * nn;
        case GRNS_OP_SCALE2POW: return c * (8 * nn + 4);
        case GRNS_OP_MAX: return c * 4 * nn;
        default: return 0;
    }
}

struct Request {
    long long ticket;
    grns_ctx *ctx;
    std::vector<grns_batch_task> tasks;
    size_t bytes;
    Clock::time_point t0;
};

struct TicketRec {
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    double latency_ms = 0;
    grns_status result = GRNS_OK;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    BatchTask *d_tasks = nullptr;
    BatchTask *h_stage = nullptr;
    unsigned long long *d_counters = nullptr;
    void *max_ws = nullptr;
    size_t max_ws_bytes = 0;
    int cap = 0;
    bool busy = false;
    std::vector<Request> reqs;
    size_t bytes = 0;
    Clock::time_point sim_done;
};

}  // namespace

struct grns_server {
    int device = 0;
    grns_serve_config cfg{};
    int cus = 256;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> pending;
    std::deque<TicketRec> tickets;
    size_t budget_free = 0;
    bool stop = false;
    std::thread worker;
    std::vector<double> lat;
    long long submitted = 0, completed = 0, rejected = 0, launches = 0;
    int sim_delay_us = -1;  // >= 0: host simulation of launches (scheduler tests without a GPU)
};

namespace {

// Tasks at least this long run as their own streaming launches (the per-call kernels) on the
// slot's stream; shorter ones are fused into one persistent launch.
constexpr size_t DIRECT_MIN = (size_t)1 << 15;

grns_status launch_direct(grns_ctx *c, const grns_batch_task &t, Slot &sl) {
    switch (t.op) {
        case GRNS_OP_EVAL:
        case GRNS_OP_EVAL_FAST:
            return grns_eval(c, t.a, t.a_mod_stride, t.a_num_stride, t.count, (grns_interval *)t.out0,
                             t.op == GRNS_OP_EVAL_FAST, sl.stream);
        case GRNS_OP_CMP:
            return grns_cmp(c, t.a, t.a_mod_stride, t.a_num_stride, t.b, t.b_mod_stride, t.b_num_stride, t.count,
                            (int *)t.out0, sl.stream);
        case GRNS_OP_MPINT_ADD: return grns_mpint_add(c, &t.mx, &t.my, &t.mz, t.count, t.overflow, sl.stream);
        case GRNS_OP_MPINT_SUB: return grns_mpint_sub(c, &t.mx, &t.my, &t.mz, t.count, t.overflow, sl.stream);
        case GRNS_OP_MPINT_MUL: return grns_mpint_mul(c, &t.mx, &t.my, &t.mz, t.count, t.overflow, sl.stream);
        case GRNS_OP_DIV:
        case GRNS_OP_DIV_FAST:
            return grns_div(c, t.a, t.b, t.a_mod_stride, t.a_num_stride, t.count, (int *)t.out0, (int *)t.out1,
                            t.op == GRNS_OP_DIV_FAST, sl.stream);
        case GRNS_OP_SCALE2POW:
            return grns_scale2pow(c, t.a, t.a_mod_stride, t.a_num_stride, t.D, t.count, (int *)t.out0, sl.stream);
        case GRNS_OP_MAX:
            return grns_max(c, t.a, t.a_mod_stride, t.a_num_stride, t.count, (int *)t.out0, (long long *)t.out1,
                            sl.max_ws, sl.max_ws_bytes, sl.stream);
        default: return GRNS_ERR_ARG;
    }
}

grns_status launch_group(grns_server *S, Slot &sl) {
    if (S->sim_delay_us >= 0) {  // host simulation: the launch "completes" after a fixed delay
        sl.sim_done = Clock::now() + std::chrono::microseconds(S->sim_delay_us);
        return GRNS_OK;
    }
    std::vector<grns_batch_task> flat, direct;
    grns_ctx *c = sl.reqs.front().ctx;
    for (auto &r : sl.reqs)
        for (auto &t : r.tasks) {
            if (t.count == 0) continue;
            if (t.op == GRNS_OP_MAX || t.count >= DIRECT_MIN) direct.push_back(t);
            else flat.push_back(t);
        }
    grns_batch_config bc{S->cfg.tile, S->cfg.blocks_per_cu, 1, 0};
    TaskPlan plan;
    make_plan(flat.data(), (int)flat.size(), c->n, bc, plan);
    grns_status st = GRNS_OK;
    if (!plan.all.empty()) {
        std::memcpy(sl.h_stage, plan.all.data(), sizeof(BatchTask) * plan.all.size());
        if (hipMemcpyAsync(sl.d_tasks, sl.h_stage, sizeof(BatchTask) * plan.all.size(), hipMemcpyHostToDevice,
                           sl.stream) != hipSuccess)
            st = GRNS_ERR_HIP;
        if (st == GRNS_OK) st = enqueue_plan(c, plan, sl.d_tasks, sl.d_counters, bc, S->cus, sl.stream);
    }
    for (auto &t : direct) {
        if (st != GRNS_OK) break;
        st = launch_direct(c, t, sl);
    }
    if (hipEventRecord(sl.done, sl.stream) != hipSuccess && st == GRNS_OK) st = GRNS_ERR_HIP;
    return st;
}

void serve_loop(grns_server *S) {
    const bool sim = S->sim_delay_us >= 0;
    if (!sim) (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        bool progress = false;
        // completions
        for (auto &sl : S->slots) {
            if (!sl.busy) continue;
            lk.unlock();
            hipError_t q = sim ? (Clock::now() >= sl.sim_done ? hipSuccess : hipErrorNotReady) : hipEventQuery(sl.done);
            lk.lock();
            if (q == hipErrorNotReady) continue;
            auto now = Clock::now();
            for (auto &r : sl.reqs) {
                TicketRec &t = S->tickets[(size_t)r.ticket];
                t.latency_ms = std::chrono::duration<double, std::milli>(now - r.t0).count();
                t.result = (q == hipSuccess) ? GRNS_OK : GRNS_ERR_HIP;
                t.state = (q == hipSuccess) ? 2 : 3;
                if (q == hipSuccess) S->lat.push_back(t.latency_ms);
                S->completed++;
            }
            S->budget_free += sl.bytes;
            sl.busy = false;
            sl.reqs.clear();
            sl.bytes = 0;
            progress = true;
        }
        if (progress) S->cv.notify_all();
        // admission: arrival order; a launch fuses pending requests of the head's moduli set that
        // fit the remaining budget
        for (auto &sl : S->slots) {
            if (sl.busy || S->pending.empty()) continue;
            if (S->pending.front().bytes > S->budget_free) break;
            grns_ctx *cx = S->pending.front().ctx;
            size_t bytes = 0;
            int ntasks = 0;
            for (auto it = S->pending.begin(); it != S->pending.end() && (int)sl.reqs.size() < S->cfg.max_batch;) {
                bool fits = it->ctx == cx && bytes + it->bytes <= S->budget_free &&
                            ntasks + (int)it->tasks.size() <= sl.cap;
                if (fits) {
                    bytes += it->bytes;
                    ntasks += (int)it->tasks.size();
                    S->tickets[(size_t)it->ticket].state = 1;
                    sl.reqs.push_back(std::move(*it));
                    it = S->pending.erase(it);
                } else {
                    ++it;
                }
            }
            S->budget_free -= bytes;
            sl.bytes = bytes;
            sl.busy = true;
            S->launches++;
            lk.unlock();
            grns_status st = launch_group(S, sl);
            lk.lock();
            if (st != GRNS_OK) {
                for (auto &r : sl.reqs) {
                    TicketRec &t = S->tickets[(size_t)r.ticket];
                    t.state = 3;
                    t.result = st;
                    S->completed++;
                }
                if (!sim) (void)hipStreamSynchronize(sl.stream);
                S->budget_free += sl.bytes;
                sl.busy = false;
                sl.reqs.clear();
                sl.bytes = 0;
                S->cv.notify_all();
            }
            progress = true;
        }
        bool idle = S->pending.empty();
        for (auto &sl : S->slots) idle = idle && !sl.busy;
        if (S->stop && idle) break;
        if (!progress) {
            if (idle) S->cv.wait(lk, [S] { return S->stop || !S->pending.empty(); });
            else S->cv.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

}  // namespace

extern "C" {

grns_status grns_serve_create(int device, const grns_serve_config *cfg, grns_server **out) {
    if (!cfg || !out || cfg->slots < 1 || cfg->slots > 16 || cfg->max_batch < 1 || cfg->hbm_budget == 0 ||
        cfg->tile < 1 || cfg->blocks_per_cu < 1 || device < 0)
        return GRNS_ERR_ARG;
    grns_server *S = new (std::nothrow) grns_server();
    if (!S) return GRNS_ERR_OOM;
    S->device = device;
    S->cfg = *cfg;
    S->budget_free = cfg->hbm_budget;
    (void)hipSetDevice(device);
    if (hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) S->cus = 256;
    S->slots.resize((size_t)cfg->slots);
    for (auto &sl : S->slots) {
        sl.cap = 8 * cfg->max_batch;
        // the largest arg-max workspace of any context (64 moduli, the most stage-1 blocks)
        sl.max_ws_bytes = (size_t)FBLOCK * MAX_FINAL_PER_THREAD * (sizeof(MaxCand) + 64 * sizeof(int));
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc(&sl.d_tasks, sizeof(BatchTask) * sl.cap) != hipSuccess ||
            hipHostMalloc(&sl.h_stage, sizeof(BatchTask) * sl.cap) != hipSuccess ||
            hipMalloc(&sl.d_counters, BATCH_COUNTERS * sizeof(unsigned long long)) != hipSuccess ||
            hipMalloc(&sl.max_ws, sl.max_ws_bytes) != hipSuccess) {
            grns_serve_destroy(S);
            return GRNS_ERR_OOM;
        }
    }
    S->worker = std::thread(serve_loop, S);
    *out = S;
    return GRNS_OK;
}

/* Scheduler test hook: a server whose launches run nothing and complete delay_us after they are
 * issued; it takes host-only contexts (device -1). */
grns_status grns_serve_create_sim(const grns_serve_config *cfg, int delay_us, grns_server **out) {
    if (!cfg || !out || cfg->slots < 1 || cfg->slots > 16 || cfg->max_batch < 1 || cfg->hbm_budget == 0 ||
        delay_us < 0)
        return GRNS_ERR_ARG;
    grns_server *S = new (std::nothrow) grns_server();
    if (!S) return GRNS_ERR_OOM;
    S->device = -1;
    S->cfg = *cfg;
    S->budget_free = cfg->hbm_budget;
    S->sim_delay_us = delay_us;
    S->slots.resize((size_t)cfg->slots);
    for (auto &sl : S->slots) sl.cap = 8 * cfg->max_batch;
    S->worker = std::thread(serve_loop, S);
    *out = S;
    return GRNS_OK;
}

grns_status grns_serve_submit(grns_server *S, grns_ctx *ctx, const grns_batch_task *tasks, int ntasks,
                              long long *ticket) {
    if (!S || !ctx || !tasks || ntasks < 1 || !ticket || ctx->device != S->device || ctx->n > 64) return GRNS_ERR_ARG;
    Request r;
    r.ctx = ctx;
    r.bytes = 0;
    for (int i = 0; i < ntasks; i++) {
        const grns_batch_task &u = tasks[i];
        if (!task_valid(u, true)) return GRNS_ERR_ARG;
        r.bytes += task_bytes(u, ctx->n);
    }
    if (ntasks > 8 * S->cfg.max_batch) return GRNS_ERR_ARG;
    r.tasks.assign(tasks, tasks + ntasks);
    std::lock_guard<std::mutex> g(S->mu);
    r.ticket = (long long)S->tickets.size();
    S->tickets.emplace_back();
    *ticket = r.ticket;
    S->submitted++;
    if (r.bytes > S->cfg.hbm_budget) {
        S->tickets.back().state = 3;
        S->tickets.back().result = GRNS_ERR_ADMISSION;
        S->rejected++;
        return GRNS_ERR_ADMISSION;
    }
    r.t0 = Clock::now();
    S->pending.push_back(std::move(r));
    S->cv.notify_all();
    return GRNS_OK;
}

grns_status grns_serve_poll(grns_server *S, long long ticket, int *state, double *latency_ms) {
    if (!S || !state) return GRNS_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    if (ticket < 0 || ticket >= (long long)S->tickets.size()) return GRNS_ERR_ARG;
    const TicketRec &t = S->tickets[(size_t)ticket];
    *state = t.state;
    if (latency_ms) *latency_ms = t.latency_ms;
    return t.state == 3 ? t.result : GRNS_OK;
}

grns_status grns_serve_wait(grns_server *S, long long ticket, double *latency_ms) {
    if (!S) return GRNS_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    if (ticket < 0 || ticket >= (long long)S->tickets.size()) return GRNS_ERR_ARG;
    S->cv.wait(lk, [&] { return S->tickets[(size_t)ticket].state >= 2; });
    const TicketRec &t = S->tickets[(size_t)ticket];
    if (latency_ms) *latency_ms = t.latency_ms;
    return t.state == 2 ? GRNS_OK : t.result;
}

grns_status grns_serve_drain(grns_server *S) {
    if (!S) return GRNS_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv.wait(lk, [S] {
        if (!S->pending.empty()) return false;
        for (auto &sl : S->slots)
            if (sl.busy) return false;
        return true;
    });
    return GRNS_OK;
}

grns_status grns_serve_stats_get(grns_server *S, grns_serve_stats *o) {
    if (!S || !o) return GRNS_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    o->submitted = S->submitted;
    o->completed = S->completed;
    o->rejected = S->rejected;
    o->launches = S->launches;
    std::vector<double> v(S->lat);
    o->p50_ms = o->p99_ms = o->mean_ms = o->max_ms = 0;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5))]; };
        o->p50_ms = pct(0.50);
        o->p99_ms = pct(0.99);
        double s = 0;
        for (double d : v) s += d;
        o->mean_ms = s / (double)v.size();
        o->max_ms = v.back();
    }
    return GRNS_OK;
}

grns_status grns_serve_destroy(grns_server *S) {
    if (!S) return GRNS_ERR_ARG;
    if (S->worker.joinable()) {
        {
            std::lock_guard<std::mutex> g(S->mu);
            S->stop = true;
        }
        S->cv.notify_all();
        S->worker.join();
    }
    for (auto &sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.d_tasks) (void)hipFree(sl.d_tasks);
        if (sl.h_stage) (void)hipHostFree(sl.h_stage);
        if (sl.d_counters) (void)hipFree(sl.d_counters);
        if (sl.max_ws) (void)hipFree(sl.max_ws);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete S;
    return GRNS_OK;
}

}  // extern "C"
