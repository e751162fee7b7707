// capi.hip -- the C ABI of include/qcmine.h.

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "qc_internal.h"

namespace qc {

// ---- device block cache (see DBuf) ----
namespace {
struct BlockCache {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void*>> free;
};
BlockCache& block_cache() {
    static BlockCache* c = new BlockCache;  // lives until the process ends
    return *c;
}
void trim_locked(BlockCache& bc, int dev) {
    for (auto& kv : bc.free)
        if (kv.first.first == dev) {
            for (void* p : kv.second) (void)hipFree(p);
            kv.second.clear();
        }
}
}  // namespace

void* pool_get(size_t cls, int dev, hipError_t* err) {
    BlockCache& bc = block_cache();
    {
        std::lock_guard<std::mutex> lock(bc.mu);
        auto it = bc.free.find({dev, cls});
        if (it != bc.free.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            *err = hipSuccess;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        // give the cached blocks of this device back and try once more
        (void)hipGetLastError();
        pool_trim(dev);
        p = nullptr;
        e = hipMalloc(&p, cls);
    }
    *err = e;
    return e == hipSuccess ? p : nullptr;
}

void pool_put(void* p, size_t cls, int dev, hipStream_t s) {
    // the block is free once the work its stream enqueued on it has ended
    // (usually already the case: owners release after reading results back)
    if (s) (void)hipStreamSynchronize(s);
    BlockCache& bc = block_cache();
    std::lock_guard<std::mutex> lock(bc.mu);
    bc.free[{dev, cls}].push_back(p);
}

void pool_trim(int dev) {
    BlockCache& bc = block_cache();
    std::lock_guard<std::mutex> lock(bc.mu);
    trim_locked(bc, dev);
}

size_t pool_cached_bytes(int dev) {
    BlockCache& bc = block_cache();
    std::lock_guard<std::mutex> lock(bc.mu);
    size_t b = 0;
    for (auto& kv : bc.free)
        if (kv.first.first == dev) b += kv.first.second * kv.second.size();
    return b;
}

int load_serialized(const char* path, qc_graph* g);
int load_adjlist(const char* path, qc_graph* g);
int write_serialized(const qc_graph* g, const char* path);
int load_binary(const char* path, qc_graph* g);
int write_binary(const qc_graph* g, const char* path);
int enumerate_persistent(const Instance* const* inst, const qc_params* params, int count, EnumOut* outs,
                         hipStream_t s);
}  // namespace qc

using namespace qc;

namespace {

// Roots and device tables of a batch, built by up to 16 host threads, each
// with a HIP stream of its own (the host filter / peel / Q-sort of one
// instance overlaps the device work of the others).
int build_instances(const qc_graph* const* graphs, const qc_params* params, int count,
                    std::vector<std::unique_ptr<Instance>>& insts) {
    int dev = 0;
    QC_HIP(hipGetDevice(&dev));
    std::vector<int> rcs(count, QC_OK);
    std::vector<std::string> errs(count);
    std::atomic<int> next{0};
    auto worker = [&]() {
        hipStream_t ts = nullptr;
        if (hipSetDevice(dev) != hipSuccess || hipStreamCreate(&ts) != hipSuccess) {
            for (int i = next++; i < count; i = next++) {
                rcs[i] = QC_EDEVICE;
                errs[i] = "cannot create a HIP stream for root construction";
            }
            return;
        }
        for (int i = next++; i < count; i = next++) {
            try {
                insts[i].reset(new Instance);
                rcs[i] = build_instance(graphs[i], params[i].gamma, params[i].min_size, *insts[i], ts);
            } catch (const std::bad_alloc&) {
                set_error("out of host memory");
                rcs[i] = QC_ENOMEM;
            }
            if (rcs[i] != QC_OK) errs[i] = last_error();
        }
        (void)hipStreamSynchronize(ts);
        (void)hipStreamDestroy(ts);
    };
    const int nthr = std::max(1, std::min(count, 16));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthr; t++) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    for (int i = 0; i < count; i++)
        if (rcs[i] != QC_OK) {
            set_error("instance %d: %s", i, errs[i].c_str());
            return rcs[i];
        }
    return QC_OK;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int check_params(const qc_params* p) {
    if (!p) {
        set_error("params is NULL");
        return QC_EINVAL;
    }
    if (!(p->gamma >= 0.5 && p->gamma <= 1.0)) {
        set_error("minimum degree ratio must be between .5 and 1 inclusive");
        return QC_EINVAL;
    }
    if (p->min_size <= 1) {
        set_error("minimum size must be greater than 1");
        return QC_EINVAL;
    }
    if (p->mode != QC_MODE_PLAIN && p->mode != QC_MODE_PERSISTENT) {
        set_error("unknown mode %d", p->mode);
        return QC_EINVAL;
    }
    if (p->order == QC_ORDER_COST && !p->cost_model) {
        set_error("cost-ordered scheduling needs a cost model");
        return QC_EINVAL;
    }
    return QC_OK;
}

int device_ok() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
        return QC_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        set_error("hipGetDeviceProperties failed");
        return QC_EDEVICE;
    }
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device is %s; this library is built for gfx950", prop.gcnArchName);
        return QC_EDEVICE;
    }
    return QC_OK;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return QC_ENOMEM;
    }
}

int finish_result(EnumOut& eo, const Instance& I, qc_result* r, double t0, double t_init, hipStream_t s) {
    int rc;
    r->raw_off = std::move(eo.raw_off);
    r->raw_v = std::move(eo.raw_v);
    r->trace_feat = std::move(eo.trace_feat);
    r->trace_parent = std::move(eo.trace_parent);
    r->trace_subtree = std::move(eo.trace_subtree);
    const double t_enum_end = eo.t_enum_end_ms > 0 ? eo.t_enum_end_ms : now_ms();
    r->st.t_init_ms = t_init - t0;
    r->st.t_enum_ms = t_enum_end - t0;
    r->st.t_kernel_ms = eo.kernel_ms;
    r->st.tasks = eo.ds.tasks;
    r->st.branches = eo.ds.branches;
    r->st.children = eo.ds.children;
    r->st.raw_cliques = eo.ds.raw_cliques;
    r->st.sum_t = eo.ds.sum_t;
    r->st.sum_t_written = eo.ds.sum_t_written;
    r->st.sum_clique_len = eo.ds.sum_clique_len;
    r->st.large_tasks = eo.ds.large_tasks;
    r->st.beyond = eo.ds.beyond;
    r->st.w_adj = eo.ds.w_adj;
    r->st.root_branches = I.root_branches;
    r->st.levels = eo.levels;
    r->st.pool_peak = eo.pool_peak;
    r->st.root_size = (uint64_t)I.R;
    if (eo.have_max) {
        r->max_off = std::move(eo.max_off);
        r->max_v = std::move(eo.max_v);
        r->st.t_nonmax_ms = eo.nonmax_ms;
    } else {
        const double tn = now_ms();
        if ((rc = remove_nonmax_device(r->raw_off, r->raw_v, r->max_off, r->max_v, s))) return rc;
        r->st.t_nonmax_ms = now_ms() - tn;
    }
    r->st.max_cliques = r->max_off.size() - 1;
    return r->st.max_cliques == 0 ? QC_EEMPTY : QC_OK;
}

}  // namespace

extern "C" {

const char* qc_last_error(void) { return qc::last_error(); }
const char* qc_version(void) { return "qcmine 0.1 (gfx950)"; }

int qc_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int qc_graph_load_serialized(const char* path, qc_graph** out) {
    if (!path || !out) return set_error("NULL argument"), QC_EINVAL;
    return guarded([&]() -> int {
        std::unique_ptr<qc_graph> g(new qc_graph);
        int rc = load_serialized(path, g.get());
        if (rc == QC_OK) *out = g.release();
        return rc;
    });
}

int qc_graph_load_adjlist(const char* path, qc_graph** out) {
    if (!path || !out) return set_error("NULL argument"), QC_EINVAL;
    return guarded([&]() -> int {
        std::unique_ptr<qc_graph> g(new qc_graph);
        int rc = load_adjlist(path, g.get());
        if (rc == QC_OK) *out = g.release();
        return rc;
    });
}

int qc_graph_write_serialized(const qc_graph* g, const char* path) {
    if (!g || !path) return set_error("NULL argument"), QC_EINVAL;
    return guarded([&]() -> int { return write_serialized(g, path); });
}

int qc_graph_write_serialized_device(const qc_graph* g, const char* path) {
    if (!g || !path) return set_error("NULL argument"), QC_EINVAL;
    int rc = device_ok();
    if (rc) return rc;
    return guarded([&]() -> int { return write_serialized_device(g, path); });
}

int qc_graph_load_binary(const char* path, qc_graph** out) {
    if (!path || !out) return set_error("NULL argument"), QC_EINVAL;
    return guarded([&]() -> int {
        std::unique_ptr<qc_graph> g(new qc_graph);
        int rc = load_binary(path, g.get());
        if (rc == QC_OK) *out = g.release();
        return rc;
    });
}

int qc_graph_write_binary(const qc_graph* g, const char* path) {
    if (!g || !path) return set_error("NULL argument"), QC_EINVAL;
    return write_binary(g, path);
}

int qc_graph_from_csr(int32_t n, const uint64_t* offsets, const uint32_t* adj, qc_graph** out) {
    if (n < 0 || !offsets || !out || (!adj && offsets[n] != 0)) return set_error("bad CSR arguments"), QC_EINVAL;
    return guarded([&]() -> int {
        std::unique_ptr<qc_graph> g(new qc_graph);
        g->n = n;
        g->off.assign(offsets, offsets + n + 1);
        if (g->off[0] != 0) return set_error("offsets[0] must be 0"), QC_EINVAL;
        g->adj.assign(adj, adj + offsets[n]);
        for (int v = 0; v < n; v++) {
            if (g->off[v + 1] < g->off[v]) return set_error("offsets not monotone at %d", v), QC_EINVAL;
            for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
                uint32_t u = g->adj[j];
                if (u >= (uint32_t)n) return set_error("neighbour %u of %d out of range", u, v), QC_EINVAL;
                if ((int)u == v) return set_error("self loop at %d", v), QC_EINVAL;
                if (j > g->off[v] && u <= g->adj[j - 1])
                    return set_error("list of %d not strictly ascending", v), QC_EINVAL;
            }
        }
        // symmetry
        for (int v = 0; v < n; v++)
            for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
                uint32_t u = g->adj[j];
                if (!std::binary_search(g->adj.begin() + g->off[u], g->adj.begin() + g->off[u + 1], (uint32_t)v))
                    return set_error("edge %d-%u not symmetric", v, u), QC_EINVAL;
            }
        *out = g.release();
        return QC_OK;
    });
}

int qc_graph_generate(int32_t config, uint64_t seed, int32_t instance, qc_graph** out, double* gamma,
                      int32_t* min_size) {
    if (!out) return set_error("NULL argument"), QC_EINVAL;
    return guarded([&]() -> int {
        std::unique_ptr<qc_graph> g(new qc_graph);
        double gm = 0;
        int ms = 0;
        int rc = generate(config, seed, instance, g.get(), &gm, &ms);
        if (rc) return rc;
        if (gamma) *gamma = gm;
        if (min_size) *min_size = ms;
        *out = g.release();
        return QC_OK;
    });
}

int qc_graph_csr(const qc_graph* g, int32_t* n, const uint64_t** offsets, const uint32_t** adj) {
    if (!g) return set_error("NULL graph"), QC_EINVAL;
    if (n) *n = g->n;
    if (offsets) *offsets = g->off.data();
    if (adj) *adj = g->adj.data();
    return QC_OK;
}

int qc_graph_planted(const qc_graph* g, uint64_t* count, const uint64_t** offsets, const uint32_t** verts) {
    if (!g) return set_error("NULL graph"), QC_EINVAL;
    if (count) *count = g->planted_off.size() - 1;
    if (offsets) *offsets = g->planted_off.data();
    if (verts) *verts = g->planted.data();
    return QC_OK;
}

uint64_t qc_graph_checksum(const qc_graph* g) { return g ? fnv1a_graph(g) : 0; }

void qc_graph_free(qc_graph* g) { delete g; }

int qc_mine(const qc_graph* g, const qc_params* p, qc_result** out) {
    if (!g || !out) return set_error("NULL argument"), QC_EINVAL;
    int rc = check_params(p);
    if (rc) return rc;
    if ((rc = device_ok())) return rc;
    return guarded([&]() -> int {
        int rc2;
        hipStream_t s;
        QC_HIP(hipStreamCreate(&s));
        std::unique_ptr<qc_result> r(new qc_result);
        // the ENUMERATION window starts with the graph resident in HBM
        // (allocate_memory precedes the timer in the reference, main.cu:545-551)
        if ((rc2 = upload_graph(g, s))) {
            (void)hipStreamDestroy(s);
            return rc2;
        }
        QC_HIP(hipDeviceSynchronize());
        double t0 = now_ms();
        Instance I;
        EnumOut eo;
        rc2 = build_instance(g, p->gamma, p->min_size, I, s);
        double t_init = now_ms();
        if (rc2 == QC_OK) {
            if (p->mode == QC_MODE_PLAIN) {
                rc2 = enumerate_plain(I, *p, eo, s);
            } else {
                const Instance* ip = &I;
                rc2 = enumerate_persistent(&ip, p, 1, &eo, s);
            }
        }
        if (rc2 == QC_OK) rc2 = finish_result(eo, I, r.get(), t0, t_init, s);
        (void)hipStreamDestroy(s);
        if (rc2 == QC_OK || rc2 == QC_EEMPTY) *out = r.release();
        return rc2;
    });
}

int qc_mine_batch(const qc_graph* const* graphs, const qc_params* params, int32_t count, qc_result** out) {
    if (!graphs || !params || !out || count < 0) return set_error("NULL argument"), QC_EINVAL;
    for (int i = 0; i < count; i++) {
        int rc = check_params(&params[i]);
        if (rc) return rc;
        if (params[i].mode != QC_MODE_PERSISTENT) {
            set_error("batched mining runs in persistent mode");
            return QC_EINVAL;
        }
    }
    int rc = device_ok();
    if (rc) return rc;
    return guarded([&]() -> int {
        int rc2 = QC_OK;
        hipStream_t s;
        QC_HIP(hipStreamCreate(&s));
        for (int i = 0; i < count && rc2 == QC_OK; i++) rc2 = upload_graph(graphs[i], s);
        QC_HIP(hipDeviceSynchronize());
        double t0 = now_ms();
        std::vector<std::unique_ptr<Instance>> insts(count);
        std::vector<const Instance*> ip(count);
        if (rc2 == QC_OK) rc2 = build_instances(graphs, params, count, insts);
        for (int i = 0; i < count; i++) ip[i] = insts[i].get();
        double t_init = now_ms();
        std::vector<EnumOut> eo(count);
        if (rc2 == QC_OK) rc2 = enumerate_persistent(ip.data(), params, count, eo.data(), s);
        std::vector<std::unique_ptr<qc_result>> rs(count);
        for (int i = 0; i < count && rc2 == QC_OK; i++) {
            rs[i].reset(new qc_result);
            int r = finish_result(eo[i], *insts[i], rs[i].get(), t0, t_init, s);
            if (r != QC_OK && r != QC_EEMPTY) rc2 = r;
        }
        (void)hipStreamDestroy(s);
        if (rc2 == QC_OK)
            for (int i = 0; i < count; i++) out[i] = rs[i].release();
        return rc2;
    });
}

int qc_result_view(const qc_result* r, uint64_t* count, const uint64_t** offsets, const uint32_t** verts) {
    if (!r) return set_error("NULL result"), QC_EINVAL;
    if (count) *count = r->max_off.size() - 1;
    if (offsets) *offsets = r->max_off.data();
    if (verts) *verts = r->max_v.data();
    return QC_OK;
}

int qc_result_raw(const qc_result* r, uint64_t* count, const uint64_t** offsets, const uint32_t** verts) {
    if (!r) return set_error("NULL result"), QC_EINVAL;
    if (count) *count = r->raw_off.size() - 1;
    if (offsets) *offsets = r->raw_off.data();
    if (verts) *verts = r->raw_v.data();
    return QC_OK;
}

int qc_result_trace(const qc_result* r, uint64_t* rows, const float** feats, const int64_t** parent,
                    const int64_t** subtree) {
    if (!r) return set_error("NULL result"), QC_EINVAL;
    if (rows) *rows = r->trace_parent.size();
    if (feats) *feats = r->trace_feat.data();
    if (parent) *parent = r->trace_parent.data();
    if (subtree) *subtree = r->trace_subtree.data();
    return QC_OK;
}

int qc_result_stats(const qc_result* r, qc_stats* s) {
    if (!r || !s) return set_error("NULL argument"), QC_EINVAL;
    *s = r->st;
    return QC_OK;
}

static int write_sets(const std::vector<uint64_t>& off, const std::vector<uint32_t>& v, const char* path) {
    FILE* f = fopen(path, "w");
    if (!f) {
        set_error("Error: cannot open file %s for write", path);
        return QC_EIO;
    }
    std::vector<char> buf;
    buf.reserve(1 << 20);
    char tmp[16];
    for (size_t i = 0; i + 1 < off.size(); i++) {
        int l = snprintf(tmp, sizeof(tmp), "%d ", (int)(off[i + 1] - off[i]));
        buf.insert(buf.end(), tmp, tmp + l);
        for (uint64_t j = off[i]; j < off[i + 1]; j++) {
            l = snprintf(tmp, sizeof(tmp), "%d ", (int)v[j]);
            buf.insert(buf.end(), tmp, tmp + l);
        }
        buf.push_back('\n');
        if (buf.size() > (1 << 20)) {
            fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    fwrite(buf.data(), 1, buf.size(), f);
    if (fclose(f) != 0) {
        set_error("write to %s failed", path);
        return QC_EIO;
    }
    return QC_OK;
}

int qc_result_write(const qc_result* r, const char* path) {
    if (!r || !path) return set_error("NULL argument"), QC_EINVAL;
    if (r->max_off.size() <= 1) {
        set_error("no quasi-cliques found; nothing written");
        return QC_EEMPTY;
    }
    return write_sets(r->max_off, r->max_v, path);
}

int qc_result_write_raw(const qc_result* r, const char* path) {
    if (!r || !path) return set_error("NULL argument"), QC_EINVAL;
    return write_sets(r->raw_off, r->raw_v, path);
}

void qc_result_free(qc_result* r) { delete r; }

int qc_degree_tables(double gamma, int32_t n_mindeg, int32_t* mindeg, int32_t n_ubq, int32_t* ubq) {
    if (n_mindeg < 0 || n_ubq < 0 || (n_mindeg && !mindeg) || (n_ubq && !ubq)) return set_error("bad arguments"), QC_EINVAL;
    std::vector<int32_t> md, uq;
    int n = std::max(n_mindeg, n_ubq);
    build_tables(gamma, n, 2, md, uq);
    for (int i = 0; i < n_mindeg; i++) mindeg[i] = md[i];
    for (int i = 0; i < n_ubq; i++) ubq[i] = uq[i];
    return QC_OK;
}

int qc_remove_nonmax(uint64_t count, const uint64_t* offsets, const uint32_t* verts, qc_result** out) {
    if (!offsets || !out || (!verts && offsets[count] != 0)) return set_error("NULL argument"), QC_EINVAL;
    int rc = device_ok();
    if (rc) return rc;
    return guarded([&]() -> int {
        hipStream_t s;
        QC_HIP(hipStreamCreate(&s));
        std::unique_ptr<qc_result> r(new qc_result);
        r->raw_off.assign(offsets, offsets + count + 1);
        r->raw_v.assign(verts, verts + offsets[count]);
        double tn = now_ms();
        int rc2 = remove_nonmax_device(r->raw_off, r->raw_v, r->max_off, r->max_v, s);
        r->st.t_nonmax_ms = now_ms() - tn;
        r->st.raw_cliques = count;
        r->st.max_cliques = r->max_off.size() - 1;
        (void)hipStreamDestroy(s);
        if (rc2) return rc2;
        qc_result* rr = r.release();
        *out = rr;
        return rr->st.max_cliques == 0 ? QC_EEMPTY : QC_OK;
    });
}

}  // extern "C"
