// persist.hip -- persistent-kernel enumeration: one launch for one or many
// mining instances.
//
// The roots of all instances seed one device FIFO queue; every wave of
// expand_persistent pops task refs (tagged with their instance), expands
// them against that instance's tables and pushes the children it writes,
// until the pending-task counter drops to zero.  There is no host round trip
// between levels (the reference returns to the host after every level,
// main.cu:606-678).  The reset of the queue state and the launch are captured
// in a hipGraph; the reset is a kernel (no memset nodes), so the graph can be
// replayed.

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "qc_internal.h"

namespace qcdev {
template <int TL>
__global__ void expand_persistent(const ExpandArgs*);

// Queue state for a fresh run: the roots dealt over the top bucket's ns
// rings (root i at position i / ns of shard i mod ns, published), every
// other cell free for its first lap, the free lists empty, counters reset.
// ctr = heads[nr], tails[nr], pool top, hungry|pending, overflow (nr =
// nb ns rings), QSTRIDE words apart.  Root records sit at the start of the
// pool; allocated blocks start at the next unit boundary.
__global__ void persist_reset(const TaskRef* __restrict__ roots, uint32_t nroots, const Vrec* __restrict__ root_verts,
                              uint64_t root_slots, TaskRef* q_refs, unsigned long long* q_seq, uint64_t q_cap, int nb,
                              int ns, Vrec* pool, unsigned long long* ctr, unsigned long long* fl_head,
                              uint32_t n_waves, StreamCtr* clq_ctr,
                              DevStats* stats, DevStats* inst_stats, uint32_t n_inst, int* error,
                              const float* __restrict__ root_feat, float* trace_feat, long long* trace_parent,
                              unsigned int* trace_beyond) {
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t nr = (uint64_t)nb * ns;
    const uint64_t top = (uint64_t)(nb - 1) * ns;  // first ring of the top bucket
    for (uint64_t i = tid; i < nr * q_cap; i += nt) {
        const uint64_t r = i / q_cap, pos = i % q_cap;
        q_seq[i] = (r >= top && pos * ns + (r - top) < nroots) ? pos + 1 : pos;
    }
    for (uint64_t i = tid; i < nroots; i += nt) q_refs[(top + i % ns) * q_cap + i / ns] = roots[i];
    if (trace_feat) {
        // (one ring: root i at position i)
        for (uint64_t i = tid; i < (uint64_t)nroots * 12; i += nt) trace_feat[i] = root_feat[i];
        for (uint64_t i = tid; i < nroots; i += nt) trace_parent[i] = -1;
        for (uint64_t i = tid; i < q_cap; i += nt) trace_beyond[i] = 0;
    }
    for (uint64_t i = tid; i < root_slots; i += nt) pool[i] = root_verts[i];
    for (uint64_t i = tid; i < (uint64_t)POOL_NCLASS * FL_SHARDS; i += nt) fl_head[i * QSTRIDE] = (1ull << 36) - 1;
    for (uint64_t i = tid; i < n_inst; i += nt) inst_stats[i] = DevStats{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (tid == 0) {
        for (uint64_t r = 0; r < nr; r++) {
            ctr[r * QSTRIDE] = 0;  // head
            // tail: the roots dealt to this ring
            ctr[(nr + r) * QSTRIDE] = r >= top && r - top < nroots ? (nroots - (r - top) + ns - 1) / ns : 0;
        }
        ctr[2 * nr * QSTRIDE] = (root_slots + POOL_UNIT - 1) / POOL_UNIT * POOL_UNIT;  // pool top
        ctr[(2 * nr + 1) * QSTRIDE] = (uint64_t)n_waves << 32 | nroots;  // hungry: every wave, pending: the roots
        ctr[(2 * nr + 2) * QSTRIDE] = 0;                      // overflow
        clq_ctr->items = 0;
        clq_ctr->slots = 0;
        *stats = DevStats{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        *error = 0;
    }
}
}  // namespace qcdev

namespace qc {

namespace {

// Per-wave HBM slab: the root-space workspace for tasks above the mid size
// and, in the same bytes, the local-matrix workspace for tasks between the
// LDS tile and the mid size (a wave runs one task at a time).
size_t ws_bytes_host(uint32_t tcap, bool mat, uint32_t mwords) {
    return (ws_bytes(tcap, mat, mwords) + 255) & ~(size_t)255;
}

constexpr int Q_SHARDS = 8;                 // rings per bucket in the dynamic order: one per XCD
constexpr int NCTR = 2 * 4 * Q_SHARDS + 4;  // heads, tails (4 buckets x Q_SHARDS rings at most), pool top,
                                            // hungry|pending, overflow, one spare

struct RunState {
    unsigned long long ctr[NCTR];
    StreamCtr clq;
    int err;
};

// Device buffers of the persistent form, kept per device between calls (grow
// only), so a steady stream of calls allocates nothing.  Calls on one device
// serialize on the mutex.  Never destroyed: the runtime may be gone at exit.
struct Workspace {
    std::mutex mu;
    DBuf pool, q, seq, slab, clq_refs, clq_verts, ctr, clq_ctr, stats, err, args, fl, priv, pbkt;
    uint64_t pool_cap = 0, q_cap = 0, clq_slots = 0, clq_refs_cap = 0, priv_cap = 0;

    static Workspace& get(int dev) {
        static std::mutex m;
        static std::vector<Workspace*> all;
        std::lock_guard<std::mutex> g(m);
        if ((int)all.size() <= dev) all.resize(dev + 1, nullptr);
        if (!all[dev]) all[dev] = new Workspace;
        return *all[dev];
    }
};

double now_steady_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int ensure(DBuf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return QC_OK;
    return b.alloc(bytes);
}

// largest task that may use task-local matrices in the slab
constexpr int MID_MAX = 4096;
// largest pair of set masks kept in LDS (more would cost resident waves)
constexpr size_t MASK_LDS_MAX = 4096;
// LDS allocation granule of a workgroup on gfx950 (bytes), measured: single-
// wave workgroups of 32000 B run 5 per 160 KB CU, of 32256 B 4; of 20480 B
// 8, of 20736 B 7
constexpr size_t LDS_GRANULE = 1280;
}  // namespace

int enumerate_persistent(const Instance* const* insts, const qc_params* params, int count, EnumOut* outs,
                         hipStream_t s) {
    int rc;
    for (int i = 0; i < count; i++) outs[i] = EnumOut();
    if (count <= 0) return QC_OK;
    if (count > 65535) {
        set_error("at most 65535 instances per batched launch");
        return QC_EINVAL;
    }
    const qc_params& p = params[0];
    const int TL = p.lds_tile == 64 ? 64 : p.lds_tile == 256 ? 256 : 128;
    const bool cost = p.order == QC_ORDER_COST;
    const int nb = cost ? 4 : 1;
    if (cost && p.trace) {
        set_error("task traces are recorded with FIFO order only");
        return QC_EINVAL;
    }
    const bool stat = p.scheduling != 0;
    if (cost && stat) {
        set_error("static scheduling deals ring positions to the waves in turn; cost order needs dynamic scheduling");
        return QC_EINVAL;
    }
    const void* kern = TL == 64    ? (const void*)qcdev::expand_persistent<64>
                       : TL == 128 ? (const void*)qcdev::expand_persistent<128>
                                   : (const void*)qcdev::expand_persistent<256>;

    // roots and per-instance tables
    std::vector<TaskRef> roots;
    std::vector<Vrec> root_verts;
    std::vector<DevGraph> graphs(count);
    std::vector<float> root_feat;
    // a window's leading removals applied at once when they are many
    std::vector<RootCut> cuts(count);
    int maxR = 1, maxrw = 1;
    for (int i = 0; i < count; i++) {
        const Instance& I = *insts[i];
        int crc = root_window_cut(I, params[i].root_begin, params[i].root_end, cuts[i], s);
        if (crc) return crc;
        graphs[i] = I.dg;
        graphs[i].root_b0 = cuts[i].b0;
        graphs[i].root_b1 = cuts[i].b1;
        graphs[i].root_cut = cuts[i].cut;
        graphs[i].max_depth = params[i].max_depth > 0 ? params[i].max_depth : -1;
        if (I.R == 0) continue;
        maxR = std::max(maxR, I.R);
        maxrw = std::max(maxrw, I.rw);
        const std::vector<Vrec>& rrec = cuts[i].cut ? cuts[i].rec : I.root;
        TaskRef r{root_verts.size(), (uint32_t)rrec.size(), 0, (uint16_t)i, I.root_bound, 0, 0};
        roots.push_back(r);
        root_verts.insert(root_verts.end(), rrec.begin(), rrec.end());
        // root features (no members: indeg 0, every vertex a candidate)
        int me = INT32_MAX, mx = INT32_MIN;
        long long se = 0, s2 = 0;
        for (const Vrec& v : I.root) {
            me = std::min(me, v.exdeg);
            mx = std::max(mx, v.exdeg);
            se += v.exdeg;
            s2 += I.d2r[v.id];
        }
        const float R = (float)I.R;
        const float f[12] = {0.0f, R, 0.0f, 0.0f, 0.0f / R, (float)me, (float)mx, (float)se / R, 0.0f, R,
                             (float)s2 / R, R};
        root_feat.insert(root_feat.end(), f, f + 12);
    }
    if (roots.empty()) return QC_OK;

    int dev = 0, n_cu = 0, per_cu = 0;
    QC_HIP(hipGetDevice(&dev));
    QC_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    // wide tasks: per-vertex arrays in LDS up to this size (dynamic LDS)
    const int wide_req = p.lds_wide < 0 ? 0 : (p.lds_wide > 0 ? (int)p.lds_wide : 512);
    const uint32_t wide = wide_req > TL ? (uint32_t)std::min(std::min(wide_req, 1024), std::max(maxR, TL)) : 0u;
    // the two set masks of tasks above the tile go to LDS after the wide
    // arrays when they are small (rowpop reads the mask beside every row word)
    const int mid_tile = std::min(p.mid_tile > 0 ? (int)p.mid_tile : 0, MID_MAX);
    const uint32_t mid = (uint32_t)std::min(maxR, mid_tile);
    const uint32_t midw = (mid + 63) / 64;
    const size_t arr_b = wide ? ws_bytes(wide, false, 0) : (size_t)0;
    const size_t mask_b = (size_t)2 * std::max<size_t>((size_t)maxrw, midw) * 8;
    const bool mask_lds = mask_b <= MASK_LDS_MAX;
    size_t lds_b = std::max(ws_bytes(TL, true, (TL + 63) / 64), arr_b + (mask_lds ? mask_b : 0));
    QC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds_b));
    {
        // A workgroup's LDS (static + dynamic) is allocated in granules of
        // LDS_GRANULE bytes, coarser than the occupancy API assumes: near a
        // boundary the API counts one more resident wave than the CU holds
        // (160 KB: 5 workgroups of 32000 B fit, of 32256 B only 4).  Waves
        // that are not resident would only start when others leave, and the
        // static order deals positions to every wave, so the grid must not
        // count them.
        int lds_cu = 0, lds_blk = 0;
        QC_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
        QC_HIP(hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        hipFuncAttributes fa{};
        QC_HIP(hipFuncGetAttributes(&fa, kern));
        const size_t st = fa.sharedSizeBytes;
        auto alloc = [&](size_t dyn) { return (dyn + st + LDS_GRANULE - 1) / LDS_GRANULE * LDS_GRANULE; };
        per_cu = std::min(per_cu, (int)((size_t)lds_cu / alloc(lds_b)));
        if (mask_lds && per_cu > 0) {
            // the resident waves leave part of the CU's LDS unused: the
            // dynamic region grows to the most that keeps per_cu waves
            // resident, and tasks above the wide tier place arrays there
            // (the masks sit at its end)
            size_t cand = (size_t)lds_cu / (size_t)per_cu / LDS_GRANULE * LDS_GRANULE;
            cand = std::min(cand, (size_t)lds_blk);
            cand = cand > st ? (cand - st) & ~(size_t)255 : 0;
            int pc = 0;
            if (cand > lds_b) QC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, kern, 64, cand));
            if (cand > lds_b && pc >= per_cu && (size_t)lds_cu / alloc(cand) >= (size_t)per_cu) lds_b = cand;
        }
    }
    // every wave must be resident: the queue hand-off waits on producers
    uint32_t grid = (uint32_t)std::max(1, n_cu) * (uint32_t)std::max(1, std::min(per_cu, 16));
    if (p.workgroups > 0) grid = std::min<uint32_t>(grid, (uint32_t)p.workgroups);

    const bool use_mid = (int)mid > TL;
    const size_t slab_b = std::max(ws_bytes_host((uint32_t)maxR, false, (uint32_t)maxrw),
                                   use_mid ? ws_bytes_host(mid, true, midw) : (size_t)0);
    size_t freeb = 0, totalb = 0;
    QC_HIP(hipMemGetInfo(&freeb, &totalb));
    freeb += pool_cached_bytes(dev);

    Workspace& ws = Workspace::get(dev);
    std::lock_guard<std::mutex> lock(ws.mu);
    const size_t budget = std::min<size_t>((freeb + ws.pool.bytes + ws.q.bytes + ws.seq.bytes + ws.slab.bytes + ws.priv.bytes +
                                            ws.clq_refs.bytes + ws.clq_verts.bytes) / 2,
                                           (size_t)128 << 30);
    if ((size_t)grid * slab_b > budget / 4) grid = (uint32_t)std::max<size_t>(1, budget / 4 / slab_b);

    // capacities: records are recycled through the pool's free lists, the
    // rings wrap and every wave keeps its own children depth first, so the
    // live set is the search frontier, not every task ever written (with a
    // trace, positions number the trace rows: the rings then never wrap and
    // every task goes through them).  They start from what the device
    // workspace already holds and grow by rerunning the (deterministic)
    // launch when a run overflows, unless the caller fixed them
    const bool fixed_tasks = p.task_capacity != 0, fixed_clq = p.clique_capacity != 0;
    const uint64_t max_pool = std::min<uint64_t>(budget / 2, (size_t)32 << 30) / sizeof(Vrec);
    // the dynamic order spreads each bucket over Q_SHARDS rings (a wave pushes
    // to its XCD's ring and takes from it first); static order deals the
    // positions of one ring, and a trace numbers the tasks by position in one
    const bool trace = p.trace != 0;
    const int ns = stat || trace ? 1 : Q_SHARDS;
    const int nr = nb * ns;
    const uint64_t max_q = (1ull << 28) / nr;
    const uint64_t max_clq = std::min<uint64_t>(budget / 16 / 4, (uint64_t)1 << 31);
    uint64_t pool_cap = fixed_tasks ? p.task_capacity
                                    : std::max<uint64_t>(ws.pool_cap, std::min<uint64_t>(max_pool, 1ull << 26));
    pool_cap = std::max<uint64_t>(pool_cap, root_verts.size() + 1024);
    uint64_t q_cap = std::max<uint64_t>(ws.q_cap / nr, std::min<uint64_t>(max_q, (1ull << 22) / ns));
    q_cap = std::max<uint64_t>(q_cap, roots.size() + 1024);
    uint64_t clq_slots = fixed_clq ? p.clique_capacity
                                   : std::max<uint64_t>(ws.clq_slots, std::min<uint64_t>(max_clq, 1ull << 24));
    uint64_t clq_refs = std::max<uint64_t>(ws.clq_refs_cap, clq_slots / 8 + 1024);
    // private stacks: a power of two of refs per wave (the stack is a ring
    // indexed by 32-bit counters)
    const uint64_t max_priv = 1ull << 20;
    uint64_t priv_cap = std::max<uint64_t>(ws.priv_cap, 8192);
    if (const char* ic = getenv("QCMINE_INITIAL_CAPACITY")) {
        // start small (tests of the grow-and-rerun path)
        const uint64_t c0 = std::max<uint64_t>(1, strtoull(ic, nullptr, 10));
        if (!fixed_tasks) pool_cap = std::max<uint64_t>(c0, root_verts.size() + 1024);
        q_cap = std::max<uint64_t>(c0, roots.size() + 1024);
        if (!fixed_clq) clq_slots = c0;
        clq_refs = c0 / 8 + 1024;
        priv_cap = 4;
        while (priv_cap * 2 <= std::min<uint64_t>(c0, 8192)) priv_cap *= 2;
    }
    // the roots' records precede the first pool block
    pool_cap = std::max<uint64_t>(pool_cap, (root_verts.size() + POOL_UNIT - 1) / POOL_UNIT * POOL_UNIT + 64);

    DBuf d_istats, d_graphs, d_roots, d_root_verts, d_tfeat, d_tpar, d_tbey, d_rfeat, d_cmw, d_cmn, d_phase, d_tl;
    // diagnostics: QCMINE_TIMELINE=<file> writes one CSV line per task taken
    const char* timeline_path = getenv("QCMINE_TIMELINE");
    if (cost && (rc = model_device_params((const qc_model*)p.cost_model, d_cmw, d_cmn, s))) return rc;
    if ((rc = ensure(ws.ctr, NCTR * QSTRIDE * 8)) || (rc = ensure(ws.fl, (size_t)POOL_NCLASS * FL_SHARDS * QSTRIDE * 8)) || (rc = ensure(ws.clq_ctr, sizeof(StreamCtr))) ||
        (rc = ensure(ws.stats, sizeof(DevStats))) || (rc = ensure(ws.err, sizeof(int))) ||
        (rc = ensure(ws.args, sizeof(ExpandArgs))) || (rc = ensure(ws.slab, (size_t)grid * slab_b)) ||
        (rc = d_istats.alloc((size_t)count * sizeof(DevStats))) ||
        (rc = d_graphs.alloc((size_t)count * sizeof(DevGraph))) ||
        (rc = d_roots.alloc(roots.size() * sizeof(TaskRef))) ||
        (rc = d_root_verts.alloc(root_verts.size() * sizeof(Vrec))))
        return rc;
    if (trace && (rc = d_rfeat.alloc(root_feat.size() * sizeof(float)))) return rc;
    QC_HIP(hipMemcpyAsync(d_graphs.p, graphs.data(), (size_t)count * sizeof(DevGraph), hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_roots.p, roots.data(), roots.size() * sizeof(TaskRef), hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_root_verts.p, root_verts.data(), root_verts.size() * sizeof(Vrec), hipMemcpyHostToDevice,
                          s));
    if (trace)
        QC_HIP(hipMemcpyAsync(d_rfeat.p, root_feat.data(), root_feat.size() * sizeof(float), hipMemcpyHostToDevice,
                              s));
    // QCMINE_PHASES=1: per-phase cycle counters, printed to stderr
    const bool phases = getenv("QCMINE_PHASES") && atoi(getenv("QCMINE_PHASES")) > 0;
    if (phases) {
        if ((rc = d_phase.alloc(PH_COUNT * 8))) return rc;
        QC_HIP(hipMemsetAsync(d_phase.p, 0, PH_COUNT * 8, s));
    }

    hipEvent_t e0, e1;
    QC_HIP(hipEventCreate(&e0));
    QC_HIP(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};

    RunState first{};
    std::vector<DevStats> first_stats;
    float first_ms = 0;
    for (int attempt = 0;; attempt++) {
        if ((rc = ensure(ws.pool, pool_cap * sizeof(Vrec))) || (rc = ensure(ws.q, nr * q_cap * sizeof(TaskRef))) ||
            (rc = ensure(ws.seq, nr * q_cap * 8)) || (rc = ensure(ws.clq_refs, clq_refs * sizeof(CliqueRef))) ||
            (rc = ensure(ws.clq_verts, clq_slots * 4)) || (rc = ensure(ws.priv, (size_t)grid * priv_cap * sizeof(TaskRef))) ||
            (rc = ensure(ws.pbkt, (size_t)grid * priv_cap)))
            return rc;
        if (trace && ((rc = d_tfeat.alloc(q_cap * 12 * sizeof(float))) || (rc = d_tpar.alloc(q_cap * 8)) ||
                      (rc = d_tbey.alloc(q_cap * 4))))
            return rc;
        ws.pool_cap = std::max(ws.pool_cap, pool_cap);
        ws.q_cap = std::max<uint64_t>(ws.q_cap, nr * q_cap);
        ws.clq_slots = std::max(ws.clq_slots, clq_slots);
        ws.clq_refs_cap = std::max(ws.clq_refs_cap, clq_refs);
        ws.priv_cap = std::max(ws.priv_cap, priv_cap);

        ExpandArgs A;
        memset(&A, 0, sizeof(A));
        A.g = graphs[0];
        A.graphs = d_graphs.as<DevGraph>();
        A.inst_stats = d_istats.as<DevStats>();
        A.in_verts = ws.pool.as<Vrec>();
        A.small_max = (uint32_t)TL;
        A.clq_refs = ws.clq_refs.as<CliqueRef>();
        A.clq_verts = ws.clq_verts.as<uint32_t>();
        A.clq_ctr = ws.clq_ctr.as<StreamCtr>();
        A.clq_ref_cap = clq_refs;
        A.clq_vert_cap = clq_slots;
        A.slab = ws.slab.as<char>();
        A.slab_bytes = slab_b;
        A.slab_tcap = (uint32_t)maxR;
        A.slab_mwords = (uint32_t)maxrw;
        A.mid_tcap = use_mid ? mid : 0;
        A.lds_wide = wide;
        A.lds_masks = mask_lds ? (int32_t)(lds_b - mask_b) : -1;
        A.stats = ws.stats.as<DevStats>();
        A.error = ws.err.as<int>();
        A.queue_mode = 1;
        A.split_chunk = p.split_chunk < 0 ? 0 : (p.split_chunk == 0 ? 64 : p.split_chunk);
        // default: split tasks above the LDS tile when one graph is mined (its
        // root and first levels are a handful of large tasks, and until they
        // are split into items a few waves work while the rest wait), above
        // 512 vertices for a batch (many roots are parallelism enough)
        const int split_min_default = count == 1 ? TL : 512;
        A.split_min_t = (uint32_t)std::max(TL, p.split_min > 0 ? (int)p.split_min : split_min_default);
        A.cost_min_t = (uint32_t)TL;
        A.q_refs = ws.q.as<TaskRef>();
        A.q_seq = ws.seq.as<unsigned long long>();
        A.fl_head = ws.fl.as<unsigned long long>();
        // magazine size: the waves' LDS stacks hold at most 2 fl_mag blocks
        // of each small class (1152 fl_mag Vrecs per wave), kept under an
        // eighth of the pool
        A.fl_mag = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(FL_LOCAL_MAX / 2, pool_cap / ((uint64_t)8 * grid * 1152)));
        A.priv = ws.priv.as<TaskRef>();
        A.priv_bkt = ws.pbkt.as<unsigned char>();
        A.priv_cap = (uint32_t)priv_cap;
        A.use_priv = trace ? 0 : 1;
        A.static_sched = stat ? 1 : 0;
        unsigned long long* ctr = ws.ctr.as<unsigned long long>();
        A.q_head = ctr;
        A.q_tail = ctr + nr * QSTRIDE;
        A.pool_top = ctr + 2 * nr * QSTRIDE;
        A.pending = ctr + (2 * nr + 1) * QSTRIDE;
        A.overflow = ctr + (2 * nr + 2) * QSTRIDE;
        A.n_buckets = nb;
        A.q_shards = ns;
        A.n_roots = (int32_t)roots.size();
        A.cm_w = cost ? d_cmw.as<uint16_t>() : nullptr;
        A.cm_norm = cost ? d_cmn.as<float>() : nullptr;
        A.q_cap = q_cap;
        A.pool = ws.pool.as<Vrec>();
        A.pool_cap = pool_cap;
        A.trace_feat = trace ? d_tfeat.as<float>() : nullptr;
        A.trace_parent = trace ? (long long*)d_tpar.p : nullptr;
        A.trace_beyond = trace ? d_tbey.as<unsigned int>() : nullptr;
        A.phase = phases ? d_phase.as<unsigned long long>() : nullptr;
        if (timeline_path) {
            if ((rc = d_tl.alloc(nr * q_cap * 32))) return rc;
            QC_HIP(hipMemsetAsync(d_tl.p, 0, nr * q_cap * 32, s));
            A.timeline = d_tl.as<unsigned long long>();
        }
        QC_HIP(hipMemcpyAsync(ws.args.p, &A, sizeof(A), hipMemcpyHostToDevice, s));
        QC_HIP(hipStreamSynchronize(s));

        // capture reset + expansion
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        const ExpandArgs* ap = ws.args.as<ExpandArgs>();
        void* kargs[] = {(void*)&ap};
        const uint32_t nroots = (uint32_t)roots.size();
        QC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(qcdev::persist_reset, dim3(1024), dim3(256), 0, s, d_roots.as<TaskRef>(), nroots,
                           d_root_verts.as<Vrec>(), (uint64_t)root_verts.size(), ws.q.as<TaskRef>(),
                           ws.seq.as<unsigned long long>(), q_cap, nb, ns, ws.pool.as<Vrec>(), ctr,
                           ws.fl.as<unsigned long long>(), grid,
                           ws.clq_ctr.as<StreamCtr>(), ws.stats.as<DevStats>(), d_istats.as<DevStats>(),
                           (uint32_t)count, ws.err.as<int>(), d_rfeat.as<float>(), A.trace_feat, A.trace_parent,
                           A.trace_beyond);
        hipError_t le = hipLaunchKernel(kern, dim3(grid), dim3(64), kargs, lds_b, s);
        hipError_t ce = hipStreamEndCapture(s, &graph);
        if (le != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            set_error("graph capture failed: %s", hipGetErrorString(le != hipSuccess ? le : ce));
            return QC_EDEVICE;
        }
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(graph);
            set_error("hipGraphInstantiate failed");
            return QC_EDEVICE;
        }
        struct GraphGuard {
            hipGraph_t g;
            hipGraphExec_t e;
            ~GraphGuard() {
                (void)hipGraphExecDestroy(e);
                (void)hipGraphDestroy(g);
            }
        } gg{graph, exec};

        const int replays = std::max(1, p.replays);
        int status = QC_OK;
        bool overflow = false;
        for (int rep = 0; rep < replays && status == QC_OK; rep++) {
            QC_HIP(hipEventRecord(e0, s));
            QC_HIP(hipGraphLaunch(exec, s));
            QC_HIP(hipEventRecord(e1, s));
            QC_HIP(hipEventSynchronize(e1));
            float ms = 0;
            QC_HIP(hipEventElapsedTime(&ms, e0, e1));
            RunState rs{};
            std::vector<DevStats> ist(count);
            {
                unsigned long long raw[NCTR * QSTRIDE];
                QC_HIP(hipMemcpy(raw, ctr, sizeof(raw), hipMemcpyDeviceToHost));
                for (int k = 0; k < NCTR; k++) rs.ctr[k] = raw[k * QSTRIDE];
            }
            QC_HIP(hipMemcpy(&rs.clq, ws.clq_ctr.p, sizeof(rs.clq), hipMemcpyDeviceToHost));
            QC_HIP(hipMemcpy(&rs.err, ws.err.p, sizeof(int), hipMemcpyDeviceToHost));
            QC_HIP(hipMemcpy(ist.data(), d_istats.p, (size_t)count * sizeof(DevStats), hipMemcpyDeviceToHost));
            if (rs.err) {
                if (rs.err == 1 && rep == 0) {
                    // grow what ran out (its counter kept counting) and rerun
                    const uint64_t pool_need = rs.ctr[2 * nr];
                    const uint64_t pushes_lost = rs.ctr[2 * nr + 2];
                    bool grown = false;
                    if (pool_need > pool_cap && !fixed_tasks && pool_cap < max_pool) {
                        pool_cap = std::min<uint64_t>(max_pool, std::max<uint64_t>(pool_cap * 4, pool_need * 2));
                        grown = true;
                    }
                    // a push found its ring and this wave's stack full (with
                    // a trace: the ring, which does not wrap)
                    if (pushes_lost > 0) {
                        if (q_cap < max_q) {
                            q_cap = std::min<uint64_t>(max_q, q_cap * 4);
                            grown = true;
                        }
                        if (!trace && priv_cap < max_priv) {
                            priv_cap = std::min<uint64_t>(max_priv, priv_cap * 4);
                            grown = true;
                        }
                    }
                    if ((rs.clq.slots > clq_slots || rs.clq.items > clq_refs) && !fixed_clq && clq_slots < max_clq) {
                        clq_slots = std::min<uint64_t>(max_clq, std::max<uint64_t>(clq_slots * 4, rs.clq.slots * 2));
                        clq_refs = std::max<uint64_t>(clq_refs * 4, rs.clq.items * 2);
                        grown = true;
                    }
                    if (grown) {
                        overflow = true;
                        break;
                    }
                }
                set_error("persistent expansion failed (error flags %d: 1 capacity, 2 task too large, 4 queue "
                          "hand-off)",
                          rs.err);
                status = (rs.err & 4) ? QC_EDEVICE : QC_EOVERFLOW;
                break;
            }
            if (rep == 0) {
                first = rs;
                first_stats = ist;
                first_ms = ms;
            } else if (first.clq.items != rs.clq.items || first.clq.slots != rs.clq.slots ||
                       memcmp(first_stats.data(), ist.data(), (size_t)count * sizeof(DevStats)) != 0) {
                set_error("graph replay %d differs from the first launch", rep);
                status = QC_EDEVICE;
            }
        }
        if (status != QC_OK) return status;
        if (overflow && attempt < 8) continue;
        if (overflow) {
            set_error("persistent expansion: capacities still exceeded after growing");
            return QC_EOVERFLOW;
        }
        break;
    }
    // the ENUMERATION window ends with the expansion done and its counters
    // read; the reference stops its timer there too, before it copies the
    // device clique buffer out (main.cu:683-688)
    const double t_enum_end = now_steady_ms();
    for (int i = 0; i < count; i++) outs[i].t_enum_end_ms = t_enum_end;
    if (timeline_path) {
        std::vector<unsigned long long> tl((size_t)nr * q_cap * 4);
        QC_HIP(hipMemcpy(tl.data(), d_tl.p, tl.size() * 8, hipMemcpyDeviceToHost));
        int khz = 0;
        (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
        if (FILE* f = fopen(timeline_path, "w")) {
            fprintf(f, "slot,start_us,end_us,wave,t,depth,b0,b1\n");
            unsigned long long t0 = ~0ull;
            for (size_t i = 0; i < tl.size() / 4; i++)
                if (tl[4 * i + 1]) t0 = std::min(t0, tl[4 * i]);
            const double us = khz > 0 ? 1000.0 / khz : 0.01;
            for (size_t i = 0; i < tl.size() / 4; i++) {
                const unsigned long long* r = &tl[4 * i];
                if (!r[1]) continue;
                fprintf(f, "%zu,%.3f,%.3f,%llu,%llu,%llu,%llu,%llu\n", i, (r[0] - t0) * us, (r[1] - t0) * us,
                        r[2] >> 32, r[2] & 0xffffffffull, r[3] & 0xffffull, (r[3] >> 16) & 0xffffffull, r[3] >> 40);
            }
            fclose(f);
        }
    }
    if (phases) {
        unsigned long long ph[PH_COUNT];
        QC_HIP(hipMemcpy(ph, d_phase.p, sizeof(ph), hipMemcpyDeviceToHost));
        static const char* names[] = {"load", "lookahead", "remove", "add", "prune", "child", "prepare", "idle",
                                      "lu", "degupd", "classify", "critical", "compact", "hist", "pool", "pop", "give", "sync"};
        double tot = 0;
        for (int k = 0; k < PH_CYCLES; k++) tot += (double)ph[k];
        fprintf(stderr, "[qcmine] phases (grid %u waves, %.3g wave-cycles):", grid, tot);
        for (int k = 0; k < PH_CYCLES; k++)
            fprintf(stderr, " %s %.1f%%", names[k], 100.0 * ph[k] / (tot > 0 ? tot : 1));
        fprintf(stderr, "; tasks from own stacks %llu, from rings %llu, handed over %llu\n", ph[PH_N_OWN],
                ph[PH_N_RING], ph[PH_N_GIVEN]);
    }

    // per-instance outputs (the last replay's buffers; replays are identical
    // in every counter, the raw sets only in their order)
    StreamCtr cc = first.clq;
    {
        std::vector<uint64_t> go;
        std::vector<uint32_t> gv;
        if ((rc = collect_raw(ws.clq_refs.as<CliqueRef>(), ws.clq_verts.as<uint32_t>(), cc.items, cc.slots, go, gv,
                              s)))
            return rc;
        if (count == 1) {
            outs[0].raw_off = std::move(go);
            outs[0].raw_v = std::move(gv);
        } else {
            // instance i holds the next first_stats[i].raw_cliques sets
            uint64_t at = 0;
            for (int i = 0; i < count; i++) {
                const uint64_t n = first_stats[i].raw_cliques;
                if (at + n > cc.items) break;
                EnumOut& o = outs[i];
                o.raw_off.resize(n + 1);
                for (uint64_t j = 0; j <= n; j++) o.raw_off[j] = go[at + j] - go[at];
                o.raw_v.assign(gv.begin() + go[at], gv.begin() + go[at + n]);
                at += n;
            }
            if (at != cc.items) {
                set_error("raw set collection: the instances count %llu sets, the stream %llu",
                          (unsigned long long)at, (unsigned long long)cc.items);
                return QC_EDEVICE;
            }
        }
    }
    for (int i = 0; i < count; i++) {
        root_cut_counters(cuts[i], insts[i]->R, (uint32_t)TL, first_stats[i]);
        outs[i].ds = first_stats[i];
        outs[i].kernel_ms = first_ms;
        outs[i].levels = 1;
        outs[i].pool_peak = first.ctr[2 * nr];
    }
    if (getenv("QCMINE_DEBUG") && atoi(getenv("QCMINE_DEBUG")) > 0) {
        // the reference's size report (main.cu:2635-2646, 2893-2907) for the
        // one persistent launch: tasks expanded, children written, the pool's
        // high-water mark (B: records are not counted, slots are) and the raw
        // cliques, then the largest buffers
        DevStats t{};
        for (const DevStats& d : first_stats) {
            t.tasks += d.tasks;
            t.sum_t += d.sum_t;
            t.children += d.children;
            t.sum_t_written += d.sum_t_written;
            t.raw_cliques += d.raw_cliques;
            t.sum_clique_len += d.sum_clique_len;
        }
        printf("L: 1 T1: %llu %llu T2: %llu %llu B: 0 %llu C: %llu %llu\n\n", t.tasks, t.sum_t, t.children,
               t.sum_t_written, (unsigned long long)first.ctr[2 * nr], t.raw_cliques, t.sum_clique_len);
        printf("\nTASKS SIZE: %llu\nBUFFER SIZE: %llu\nCLIQUES SIZE: %llu\nCLIQUES OFFSET SIZE: %llu\n\n",
               (unsigned long long)pool_cap, (unsigned long long)first.ctr[2 * nr],
               (unsigned long long)first.clq.slots, (unsigned long long)first.clq.items);
        fflush(stdout);
    }
    // RemoveNonMax runs on the raw sets where the kernel left them, all
    // instances in one pass (instance i's vertices are keyed from base[i])
    const double t_nm0 = now_steady_ms();
    if (cc.items) {
        std::vector<uint32_t> base(count), limit(count);
        uint64_t nk = 0;
        for (int i = 0; i < count; i++) {
            base[i] = (uint32_t)nk;
            limit[i] = (uint32_t)insts[i]->n;
            nk += (uint64_t)insts[i]->n;
        }
        if (nk >= 0xFFFFFFFFull) {
            set_error("RemoveNonMax: the batch holds more than 2^32 - 1 vertices");
            return QC_EOVERFLOW;
        }
        DBuf d_base, d_limit;
        if ((rc = d_base.alloc((size_t)count * 4, s)) || (rc = d_limit.alloc((size_t)count * 4, s))) return rc;
        QC_HIP(hipMemcpyAsync(d_base.p, base.data(), (size_t)count * 4, hipMemcpyHostToDevice, s));
        QC_HIP(hipMemcpyAsync(d_limit.p, limit.data(), (size_t)count * 4, hipMemcpyHostToDevice, s));
        NmInput in;
        in.k = cc.items;
        in.nv = cc.slots;
        in.refs = ws.clq_refs.as<CliqueRef>();
        in.verts = ws.clq_verts.as<uint32_t>();
        in.base = d_base.as<uint32_t>();
        in.limit = d_limit.as<uint32_t>();
        in.nkeys = (uint32_t)std::max<uint64_t>(nk, 1);
        std::vector<std::vector<uint64_t>*> mo(count);
        std::vector<std::vector<uint32_t>*> mv(count);
        for (int i = 0; i < count; i++) {
            mo[i] = &outs[i].max_off;
            mv[i] = &outs[i].max_v;
        }
        if ((rc = remove_nonmax_sets(in, count, mo, mv, s))) return rc;
    }
    const double nm_ms = now_steady_ms() - t_nm0;
    for (int i = 0; i < count; i++) {
        outs[i].have_max = true;
        outs[i].nonmax_ms = nm_ms;
    }
    if (trace) {
        // rows = expanded tasks (split items excluded), per instance; a
        // child's slot is always above its parent's, so one backward pass
        // accumulates subtree sizes
        const uint64_t tail = std::min<uint64_t>(first.ctr[nr], q_cap);
        std::vector<float> tf(tail * 12);
        std::vector<long long> tp(tail);
        std::vector<TaskRef> qr(tail);
        std::vector<unsigned int> tb(tail);
        QC_HIP(hipMemcpy(tf.data(), d_tfeat.p, tail * 12 * sizeof(float), hipMemcpyDeviceToHost));
        QC_HIP(hipMemcpy(tp.data(), d_tpar.p, tail * 8, hipMemcpyDeviceToHost));
        QC_HIP(hipMemcpy(qr.data(), ws.q.p, tail * sizeof(TaskRef), hipMemcpyDeviceToHost));
        QC_HIP(hipMemcpy(tb.data(), d_tbey.p, tail * 4, hipMemcpyDeviceToHost));
        // records per subtree: the task, its children past the depth window
        // and its expanded children's subtrees
        std::vector<int64_t> sub(tail, 1), row(tail, -1);
        for (uint64_t i = 0; i < tail; i++) sub[i] += tb[i];
        for (uint64_t i = tail; i-- > 0;)
            if (tp[i] >= 0) sub[tp[i]] += sub[i];
        for (uint64_t i = 0; i < tail; i++) {
            if (tp[i] == -2) continue;
            EnumOut& o = outs[qr[i].inst];
            row[i] = (int64_t)o.trace_parent.size();
            o.trace_feat.insert(o.trace_feat.end(), tf.begin() + i * 12, tf.begin() + i * 12 + 12);
            o.trace_parent.push_back(tp[i] >= 0 ? row[tp[i]] : -1);
            o.trace_subtree.push_back(sub[i]);
        }
    }
    return QC_OK;
}

}  // namespace qc
