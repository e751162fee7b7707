// qc_internal.h -- library-internal declarations shared by the host drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "qc_device.h"
#include "qc_host.h"

struct qc_result {
    // maximal sets, lexicographic order, each ascending
    std::vector<uint64_t> max_off{0};
    std::vector<uint32_t> max_v;
    // raw sets as emitted
    std::vector<uint64_t> raw_off{0};
    std::vector<uint32_t> raw_v;
    qc_stats st{};
    // task trace (qc_params.trace)
    std::vector<float> trace_feat;
    std::vector<int64_t> trace_parent, trace_subtree;
};

namespace qc {

const char* last_error();

#define QC_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            qc::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return QC_EDEVICE;                                                                 \
        }                                                                                      \
    } while (0)

// Device blocks up to POOL_MAX bytes are recycled through a per-device
// cache (power-of-two classes) instead of a hipMalloc / hipFree pair per
// call; larger ones are allocated and freed directly.  A block goes back to
// the cache only once no enqueued work uses it: an owner that names the
// stream its work runs on has that stream synchronized on release; a block
// allocated without a stream is released only after its owner synchronized
// the work using it.
constexpr size_t POOL_MAX = (size_t)64 << 20;
inline size_t pool_class(size_t b) {
    if (b > POOL_MAX) return 0;
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
}
void* pool_get(size_t cls, int dev, hipError_t* err);
void pool_put(void* p, size_t cls, int dev, hipStream_t s);
// frees every cached block of `dev` (after a failed allocation)
void pool_trim(int dev);
// bytes held by the cache for `dev` (free to reuse; hipMemGetInfo misses them)
size_t pool_cached_bytes(int dev);

// owning device buffer
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    size_t cls = 0;  // pool class, 0 = not pooled
    int dev = 0;
    hipStream_t stream = nullptr;  // stream the buffer's work runs on (nullptr: synchronized by the owner)
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            if (cls)
                pool_put(p, cls, dev, stream);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        cls = 0;
        stream = nullptr;
    }
    // `s`: the stream that reads and writes the buffer; its release then
    // waits for that stream, so an early release (an error return while
    // kernels still run) is safe.  The stream must outlive the buffer.
    int alloc(size_t b, hipStream_t s = nullptr) {
        release();
        if (b == 0) b = 16;
        (void)hipGetDevice(&dev);
        const size_t c = pool_class(b);
        hipError_t e = hipSuccess;
        if (c) {
            p = pool_get(c, dev, &e);
        } else {
            e = hipMalloc(&p, b);
            if (e != hipSuccess) {
                // cached blocks are not counted as free by the runtime
                (void)hipGetLastError();
                pool_trim(dev);
                p = nullptr;
                e = hipMalloc(&p, b);
            }
        }
        if (e != hipSuccess || !p) {
            p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
            return QC_ENOMEM;
        }
        bytes = b;
        cls = c;
        stream = s;
        return QC_OK;
    }
    // exchanges the two buffers (pointer, size, pool class, device, stream)
    void swap(DBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(cls, o.cls);
        std::swap(dev, o.dev);
        std::swap(stream, o.stream);
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

// Root task and device tables of one mining instance.
struct Instance {
    int32_t n = 0;
    int32_t R = 0;
    int32_t rw = 0;
    double gamma = 0;
    int32_t min_size = 0;
    std::vector<Vrec> root;           // root task record (root-local ids)
    std::vector<uint32_t> orig;       // root-local -> input id
    std::vector<int32_t> mindeg, ubq;
    std::vector<int32_t> d2r;         // |N2(a) n root|
    uint64_t root_bound = 0;          // children slot bound of the root task
    uint64_t root_branches = 0;       // branches of the root task
    DBuf A1, A2, d_mindeg, d_ubq, d_orig, d_d2r, d_d1r;
    DevGraph dg{};
};

// uploads the graph's arrays to HBM once (root.hip)
int upload_graph(const qc_graph* g, hipStream_t s);
// the graph's CSR arrays in HBM (uploaded on first use; owned by the graph)
int graph_device_arrays(const qc_graph* g, const uint64_t** off, const uint32_t** adj, hipStream_t s);

// the serialized text file with N2 built on the GPU (n2.hip)
int write_serialized_device(const qc_graph* g, const char* path);

// builds the root task and the device tables (root.hip)
int build_instance(const qc_graph* g, double gamma, int min_size, Instance& inst, hipStream_t s);

// A window that starts at root branch b > 0 removes the root's last b
// vertices before its first branch.  The root has no members, so none of
// these removals can fail: root_window_cut applies them at once on the GPU
// (many waves; inside the expansion one wave would) and returns the shorter
// root record with the window shifted to start at 0, plus what the skipped
// steps would have added to the counters.  cut = 0: use the root as built.
struct RootCut {
    int32_t cut = 0;          // root vertices removed
    std::vector<Vrec> rec;    // the root record without them (cut > 0)
    long long b0 = 0, b1 = -1;  // the window for that record (DevGraph root_b0 / root_b1)
    uint64_t w_adj = 0;       // adjacency scans of the removals and of the lookahead on the removed vertices
};
int root_window_cut(const Instance& I, long long root_begin, long long root_end, RootCut& c, hipStream_t s);
// the counters of a launch on a cut root, completed to those of the whole
// root: its scans, its size, and whether it counts as a large task (t > small_max)
inline void root_cut_counters(const RootCut& c, int32_t R, uint32_t small_max, DevStats& d) {
    if (c.cut <= 0) return;
    d.w_adj += c.w_adj;
    d.sum_t += (unsigned long long)c.cut;
    const int32_t ts = R - c.cut;
    d.large_tasks += (unsigned long long)((uint32_t)R > small_max) - (unsigned long long)((uint32_t)ts > small_max);
}

// plain / persistent enumeration (driver.hip, persist.hip)
struct EnumOut {
    std::vector<uint64_t> raw_off{0};
    std::vector<uint32_t> raw_v;
    std::vector<float> trace_feat;
    std::vector<int64_t> trace_parent, trace_subtree;
    DevStats ds{};
    double kernel_ms = 0;
    uint64_t levels = 0;
    uint64_t pool_peak = 0;  // persistent form: pool high-water mark (vertex slots)
    // maximal sets, when the enumeration already removed the non-maximal
    // ones from its device buffers (persistent form)
    bool have_max = false;
    std::vector<uint64_t> max_off{0};
    std::vector<uint32_t> max_v;
    double t_enum_end_ms = 0;  // steady-clock ms at the end of the ENUMERATION window (0: not set)
    double nonmax_ms = 0;
};
int enumerate_plain(const Instance& inst, const qc_params& p, EnumOut& out, hipStream_t s);

// cost model weights / norms in device memory (mlp.hip)
int model_device_params(const qc_model* m, DBuf& pb, DBuf& norm, hipStream_t s);

// RemoveNonMax on the device (nonmax.hip).  Raw sets in device memory:
// `refs` (k CliqueRefs into `verts`) or, with refs == nullptr, the sets
// [offsets[i], offsets[i+1]) of `verts`.  With `base` set, a vertex v of
// instance i must be below limit[i] and is compared as the key base[i] + v
// (base, limit: device arrays); without it, vertices are below nkeys.
// nkeys bounds every key.
struct NmInput {
    uint64_t k = 0;   // sets
    uint64_t nv = 0;  // vertices over all sets
    const CliqueRef* refs = nullptr;
    const uint64_t* offsets = nullptr;
    const uint32_t* verts = nullptr;
    const uint32_t* base = nullptr;
    const uint32_t* limit = nullptr;
    uint32_t nkeys = 0;
};
// Device outputs: the count (~0 when a vertex was out of range), offsets
// (count + 1), vertices, and the instance of every set (may be nullptr).
// Capacities: k + 1 offsets, nv vertices, k instances.
struct NmOutput {
    uint64_t* count;
    uint64_t* off;
    uint32_t* v;
    uint32_t* inst;
};
// scratch bytes nonmax_run needs for k sets, nv vertices and nkeys keys
size_t nonmax_scratch_bytes(uint64_t k, uint64_t nv, uint32_t nkeys, hipStream_t s);
// enqueues the whole RemoveNonMax on `s`: no allocation, no synchronization
int nonmax_run(const NmInput& in, char* scratch, size_t scratch_bytes, const NmOutput& out, hipStream_t s);
// nonmax_run in allocated scratch, results copied to the host per instance
int remove_nonmax_sets(const NmInput& in, int n_inst, std::vector<std::vector<uint64_t>*> mo,
                       std::vector<std::vector<uint32_t>*> mv, hipStream_t s);
// the raw sets of a persistent launch's clique stream (k refs, nv vertex
// slots) as one array, instances in order and each instance's sets in
// vertex-slot order, on the host
int collect_raw(const CliqueRef* d_refs, const uint32_t* d_verts, uint64_t k, uint64_t nv, std::vector<uint64_t>& off,
                std::vector<uint32_t>& v, hipStream_t s);
// host raw sets (uploaded once)
int remove_nonmax_device(const std::vector<uint64_t>& off, const std::vector<uint32_t>& v, std::vector<uint64_t>& mo,
                         std::vector<uint32_t>& mv, hipStream_t s);

}  // namespace qc
