// qc_device.h -- device-side data layout shared by the HIP sources and the
// host drivers.
#pragma once

#include <cstdint>

// A task record in HBM: `t` vertex entries at `off` in a vertex pool.
// `bound` caps the vertex slots all children of the task can occupy:
// a child of branch vertex c holds at most min(t, members + 1 + d2(c))
// vertices, d2(c) = |N2(c) n root|.
// b1 > 0 marks a split item: branches [b0, b1) of a task whose record at
// `off` was prepared by expand (label | lvl2adj << 4, exdeg before any
// removal).
// stride (in 64-bit words) between the persistent queue's shared counters
constexpr int QSTRIDE = 32;

// Vertex pool of the persistent form: records live in blocks of whole
// units of POOL_UNIT Vrecs (256 B), in size classes of 1..8 units and then
// four steps per octave (10, 12, 14, 16, 20, 24, ...); a freed block goes on
// a free list of its class and is reused by the next record of that class.
constexpr uint32_t POOL_UNIT = 16;
constexpr int POOL_NCLASS = 112;
constexpr unsigned long long PH_HUNGRY = 1ull << 32;  // one hungry wave in ExpandArgs::pending
constexpr int FL_SHARDS = 16;     // shared free lists per class (a wave uses blockIdx % FL_SHARDS)
constexpr int FL_SMALL = 8;       // classes with per-wave LDS stacks (1..8 units)
constexpr int FL_LOCAL_MAX = 8;   // LDS stack entries per small class
__host__ __device__ inline int pool_class(uint32_t nvrec, uint32_t& units) {
    uint32_t u = (nvrec + POOL_UNIT - 1) / POOL_UNIT;
    if (u == 0) u = 1;
    if (u <= 8) {
        units = u;
        return (int)u - 1;
    }
    int e = 0;  // u - 1 in [2^e, 2^(e+1)), e >= 3
    for (uint32_t x = u - 1; x > 1; x >>= 1) e++;
    const uint32_t step = 1u << (e - 2);
    const uint32_t k = (u - (1u << e) + step - 1) / step;  // 1..4
    units = (1u << e) + k * step;
    return 8 + (e - 3) * 4 + (int)(k - 1);
}

__host__ __device__ constexpr uint32_t pow2_at_least(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Bytes of a per-wave workspace for tasks of up to `tcap` vertices: 15
// per-vertex int arrays of pow2(tcap) entries, then (`mat`) the task-local
// 1-hop / 2-hop matrices (tcap rows of ceil(tcap/64) words), then two masks of
// `mwords` words (ceil(tcap/64) with local matrices, rw in root space).
__host__ __device__ constexpr size_t ws_bytes(uint32_t tcap, bool mat, uint32_t mwords) {
    uint32_t T2 = pow2_at_least(tcap < 64 ? 64 : tcap);
    uint32_t W = (tcap + 63) / 64;
    size_t b = 0;
    b += (size_t)11 * T2 * 4;        // parent + child arrays
    b += (size_t)4 * T2 * 4;         // lists / keys
    b = (b + 15) & ~(size_t)15;
    if (mat) b += (size_t)2 * tcap * W * 8;
    b += (size_t)2 * mwords * 8;
    return b;
}

struct TaskRef {
    uint64_t off;
    uint32_t t;
    uint16_t depth;
    uint16_t inst;   // mining instance (batched launches)
    uint64_t bound;
    uint32_t b0;
    uint32_t b1;
};

// One task vertex (the reference's Vertex, main.cu:71-79, without lvl2adj,
// which every record carries as 0 and only the parent's lookahead fills).
// `id` is the root-local vertex id (0 .. R-1).
struct __attribute__((aligned(16))) Vrec {
    int32_t id;
    int32_t label;  // 1 member, 2 covered (root only), 3 cover (root only), 0 candidate
    int32_t indeg;
    int32_t exdeg;
};

// Output stream of raw cliques: `len` original vertex ids at `off`.
struct CliqueRef {
    uint64_t off;
    uint32_t len;
    uint32_t inst;
};

// Counters of one output stream (children or cliques).
struct StreamCtr {
    unsigned long long items;   // records
    unsigned long long slots;   // vertex slots
};

// Per-call statistics accumulated by the kernels.
struct DevStats {
    unsigned long long tasks, branches, children, raw_cliques, sum_t, sum_t_written, sum_clique_len, large_tasks,
        beyond, w_adj;
};

// Graph tables on the device: root-space 1-hop / 2-hop bit matrices
// (row a, bit b set iff b in N(a) / N2(a); rows of `rw` 64-bit words),
// the integer tables and the id translation.
struct DevGraph {
    int32_t R;            // root vertices
    int32_t rw;           // words per root-space row
    const uint64_t* A1;   // R x rw
    const uint64_t* A2;   // R x rw
    const int32_t* mindeg;
    int32_t mindeg_n;
    const int32_t* ubq;
    int32_t ubq_n;
    int32_t min_size;
    int32_t pad;
    const uint32_t* orig; // root-local id -> input vertex id
    const int32_t* d2r;   // |N2(a) n root| per root vertex
    const int32_t* d1r;   // |N(a) n root| per root vertex
    // search window (see qc_params)
    long long root_b0, root_b1;
    int32_t max_depth;
    // root vertices a window's leading branches removed before the launch
    // (root_window_cut): the root record then holds the rest, and its
    // lookahead, known to fail on the whole root, is not repeated; 0: none
    int32_t root_cut;
    // mindeg[i] = ceil(gamma (i - 1)) is evaluated from this double in the
    // kernels with the same expression as the host table (one correctly
    // rounded product, then ceil), so both give the same integers
    double gamma;
};

// Arguments of one expansion launch.
struct ExpandArgs {
    DevGraph g;                      // single-instance launches
    const DevGraph* graphs;          // batched launches: per-instance tables (NULL: use g)
    DevStats* inst_stats;            // batched launches: per-instance counters
    const TaskRef* in_refs;
    const Vrec* in_verts;
    uint32_t n_in;
    int32_t static_sched;
    unsigned int* task_counter;

    // children: tasks with t <= small_max go to stream 0, others to stream 1
    TaskRef* out_refs[2];
    Vrec* out_verts[2];
    StreamCtr* out_ctr;          // [2]
    unsigned long long out_ref_cap[2];
    unsigned long long out_vert_cap[2];
    uint32_t small_max;
    uint32_t pad0;

    CliqueRef* clq_refs;
    uint32_t* clq_verts;
    StreamCtr* clq_ctr;
    unsigned long long clq_ref_cap;
    unsigned long long clq_vert_cap;

    // per-wave HBM workspace for tasks above the LDS tile
    char* slab;
    unsigned long long slab_bytes;   // per wave
    uint32_t slab_tcap;              // vertices per slab
    uint32_t slab_mwords;            // mask words per slab (root-space row words)
    uint32_t mid_tcap;               // persistent mode: tasks above the LDS tile and up to
                                     // this size use task-local matrices in the slab (0: none)
    uint32_t lds_wide;               // persistent mode: tasks above the LDS tile and up to
                                     // this size keep their per-vertex arrays in LDS (0: none)
    int32_t lds_masks;               // persistent mode: byte offset in the dynamic LDS of the
                                     // two set masks of tasks above the LDS tile (-1: slab)

    // persistent mode: n_buckets x q_shards shared ring queues of task refs
    // plus one private LIFO stack per wave, over a vertex pool with free
    // lists.  Ring r = b q_shards + s (bucket b, shard s) owns cells
    // [r q_cap, (r+1) q_cap) of q_refs / q_seq and the position counters
    // q_head[r QSTRIDE], q_tail[r QSTRIDE] (every shared counter sits on a
    // 256-byte line of its own).  A wave pushes to its own shard and takes
    // from it first, then from the other shards of the same bucket.
    // Position p of a ring uses cell p mod q_cap, whose sequence word reads
    // p while the cell is free for p, p + 1 once the ref of p is published
    // and p + q_cap once it has been taken.  `pending` counts the tasks
    // waiting in the rings plus the waves holding work (0 = search
    // complete).  With a cost model, a child's bucket is its predicted log
    // subtree size; waves take the highest non-empty bucket first.
    int32_t queue_mode;
    int32_t split_chunk;             // branches per split item (0: no splitting)
    uint32_t split_min_t;            // split tasks above this size
    uint32_t cost_min_t;             // cost order: children of at most this size skip the model (bucket 0)
    TaskRef* q_refs;
    unsigned long long* q_seq;
    unsigned long long* q_head;
    unsigned long long* q_tail;
    unsigned long long q_cap;
    Vrec* pool;
    unsigned long long* pool_top;    // bump pointer for blocks no free list holds (Vrec units)
    unsigned long long pool_cap;
    unsigned long long* fl_head;     // shared stacks of free-block magazines, POOL_NCLASS x FL_SHARDS
                                     // lines of QSTRIDE words: (tag << 36) | head block index
                                     // (POOL_UNIT-Vrec units), index all ones = empty
    uint32_t fl_mag;                 // blocks per magazine of the small classes (1..FL_LOCAL_MAX / 2)
    // (hungry << 32) | pending in one word, so a wave's transition between
    // holding work and waiting is one atomic: `pending` counts the tasks in
    // the rings plus the waves holding work, `hungry` the waves holding
    // none (they ask the busy ones to hand some over).  Both stay below 2^31
    // and never go negative: every decrement follows its increment.
    unsigned long long* pending;
    unsigned long long* overflow;    // pushes that found the ring and the private stack full
    // private stacks: wave w owns priv[w priv_cap, (w+1) priv_cap) (refs)
    // and priv_bkt[...] (the bucket each ref would take in the rings);
    // children stay there (depth first) unless use_priv is 0, and the
    // oldest ones are handed to the rings while they run low
    TaskRef* priv;
    unsigned char* priv_bkt;
    uint32_t priv_cap;
    int32_t use_priv;

    // task trace (persistent mode, NULL = off): features of the task in
    // queue slot i at trace_feat[12 i], its parent's slot at trace_parent[i]
    // (-1 root, -2 split item)
    float* trace_feat;
    long long* trace_parent;
    unsigned int* trace_beyond;  // per slot: children past the depth window

    int32_t n_buckets;               // 1 = FIFO
    int32_t n_roots;                 // root tasks of the launch (root split width)
    int32_t q_shards;                // rings per bucket (ring b q_shards + s; wave w uses shard w mod q_shards)
    const uint16_t* cm_w;            // cost model: bf16 weights (5057)
    const float* cm_norm;            // mu[12] sd[12] ymu ysd thresholds[3]

    DevStats* stats;
    int* error;  // bit 0: output overflow, bit 1: task above slab capacity, bit 2: queue wait limit

    // diagnostics (NULL = off): shader cycles per phase summed over waves,
    // indexed by the PH_* constants
    unsigned long long* phase;
    // diagnostics: per queue slot {start, end, wave << 32 | t, depth | b0 << 16 | b1 << 40}
    // in wall-clock ticks, or NULL
    unsigned long long* timeline;
};

enum PhaseId {
    PH_LOAD = 0,      // record load and local matrices
    PH_LOOKAHEAD,     // lookahead
    PH_REMOVE,        // remove_one_vertex steps and item catch-up
    PH_ADD,           // add_one_vertex and diameter pruning
    PH_PRUNE,         // degree / critical pruning and the clique test
    PH_CHILD,         // child sort and write
    PH_PREPARE,       // split preparation
    PH_IDLE,          // persistent form: waiting for work (empty polls and back-off)
    PH_LU,            // LU bounds and member checks
    PH_DEGUPD,        // degree-pruning exdeg updates
    PH_COMPACT,       // degree-pruning classification
    PH_CRIT,          // critical-vertex step
    PH_MOVE,          // degree-pruning compaction of the kept candidates
    PH_HIST,          // candidate indegree tables for the LU bounds
    PH_POOL,          // persistent form: record blocks taken and given back, task refs pushed
    PH_POP,           // persistent form: taking the next task (own stack, rings)
    PH_GIVE,          // persistent form: deciding on and making hand-overs from the own stack
    PH_SYNC,          // persistent form: a ring claim's acquire and the cells' return
    PH_CYCLES,        // (the entries above count cycles, the ones below events)
    PH_N_OWN = PH_CYCLES,  // tasks taken from the wave's own stack
    PH_N_RING,        // tasks taken from the rings
    PH_N_GIVEN,       // records handed over to the rings
    PH_COUNT
};
