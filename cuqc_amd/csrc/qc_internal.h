// qc_internal.h -- library-internal declarations shared by the host drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
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
// the cache only when its owner releases it, after the work using it has
// been synchronized.
constexpr size_t POOL_MAX = (size_t)64 << 20;
inline size_t pool_class(size_t b) {
    if (b > POOL_MAX) return 0;
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
}
void* pool_get(size_t cls, int dev, hipError_t* err);
void pool_put(void* p, size_t cls, int dev);

// owning device buffer
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    size_t cls = 0;  // pool class, 0 = not pooled
    int dev = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            if (cls)
                pool_put(p, cls, dev);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        cls = 0;
    }
    int alloc(size_t b) {
        release();
        if (b == 0) b = 16;
        (void)hipGetDevice(&dev);
        const size_t c = pool_class(b);
        hipError_t e = hipSuccess;
        if (c)
            p = pool_get(c, dev, &e);
        else
            e = hipMalloc(&p, b);
        if (e != hipSuccess || !p) {
            p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
            return QC_ENOMEM;
        }
        bytes = b;
        cls = c;
        return QC_OK;
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

// builds the root task and the device tables (root.hip)
int build_instance(const qc_graph* g, double gamma, int min_size, Instance& inst, hipStream_t s);

// plain / persistent enumeration (driver.hip, persist.hip)
struct EnumOut {
    std::vector<uint64_t> raw_off{0};
    std::vector<uint32_t> raw_v;
    std::vector<float> trace_feat;
    std::vector<int64_t> trace_parent, trace_subtree;
    DevStats ds{};
    double kernel_ms = 0;
    uint64_t levels = 0;
};
int enumerate_plain(const Instance& inst, const qc_params& p, EnumOut& out, hipStream_t s);

// cost model weights / norms in device memory (mlp.hip)
int model_device_params(const qc_model* m, DBuf& pb, DBuf& norm, hipStream_t s);

// RemoveNonMax on the device (nonmax.hip)
int remove_nonmax_device(const std::vector<uint64_t>& off, const std::vector<uint32_t>& v, std::vector<uint64_t>& mo,
                         std::vector<uint32_t>& mv, hipStream_t s);

}  // namespace qc
