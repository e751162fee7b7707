// persist.hip -- persistent-kernel enumeration (one launch per call).
//
// The root record is the first entry of a device FIFO queue; every wave of
// expand_persistent pops task refs, expands them and pushes the children it
// writes, until the pending-task counter drops to zero.  There is no host
// round trip between levels (the reference returns to the host after every
// level, main.cu:606-678).

#include <algorithm>
#include <cstring>

#include "qc_internal.h"

namespace qcdev {
template <int TL>
__global__ void expand_persistent(const ExpandArgs*);
}  // namespace qcdev

namespace qc {

namespace {
size_t slab_bytes(uint32_t tcap, uint32_t rw) {
    uint32_t T2 = 64;
    while (T2 < tcap) T2 <<= 1;
    size_t b = (size_t)15 * T2 * 4;
    b = (b + 15) & ~(size_t)15;
    b += (size_t)2 * rw * 8;
    return (b + 255) & ~(size_t)255;
}
}  // namespace

int enumerate_persistent(const Instance* const* insts, const qc_params* params, int count, EnumOut* outs,
                         hipStream_t s) {
    if (count != 1) {
        set_error("batched persistent launch: not available in this build");
        return QC_EINVAL;
    }
    const Instance& I = *insts[0];
    const qc_params& p = params[0];
    EnumOut& out = outs[0];
    out = EnumOut();
    if (I.R == 0) return QC_OK;
    int rc;
    const int TL = p.lds_tile == 64 ? 64 : p.lds_tile == 256 ? 256 : 128;
    const void* kern = TL == 64    ? (const void*)qcdev::expand_persistent<64>
                       : TL == 128 ? (const void*)qcdev::expand_persistent<128>
                                   : (const void*)qcdev::expand_persistent<256>;

    int dev = 0, n_cu = 0, per_cu = 0;
    QC_HIP(hipGetDevice(&dev));
    QC_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    QC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0));
    // every wave must be resident: the queue hand-off waits on producers
    uint32_t grid = (uint32_t)std::max(1, n_cu) * (uint32_t)std::max(1, per_cu);
    if (p.workgroups > 0) grid = std::min<uint32_t>(grid, (uint32_t)p.workgroups);

    size_t freeb = 0, totalb = 0;
    QC_HIP(hipMemGetInfo(&freeb, &totalb));
    const size_t slab_b = slab_bytes((uint32_t)I.R, (uint32_t)I.rw);
    const size_t budget = std::min<size_t>(freeb / 2, (size_t)128 << 30);
    if ((size_t)grid * slab_b > budget / 4) grid = (uint32_t)std::max<size_t>(1, budget / 4 / slab_b);
    // capacities: the FIFO keeps every record it ever held (no reuse), so
    // they bound the number of tasks / vertex slots / raw cliques per call
    const uint64_t pool_cap =
        p.task_capacity ? p.task_capacity : std::min<uint64_t>(budget / 2, (size_t)32 << 30) / sizeof(Vrec);
    const uint64_t q_cap = std::min<uint64_t>(pool_cap / 8 + 1024, (uint64_t)1 << 26);
    const uint64_t clq_slots =
        p.clique_capacity ? p.clique_capacity : std::min<uint64_t>(budget / 16 / 4, (uint64_t)1 << 30);
    const uint64_t clq_refs = std::min<uint64_t>(clq_slots / 4 + 1024, (uint64_t)1 << 26);

    DBuf d_pool, d_q, d_ready, d_ctr, d_slab, d_clq_refs, d_clq_verts, d_clq_ctr, d_stats, d_err, d_args;
    if ((rc = d_pool.alloc(pool_cap * sizeof(Vrec))) || (rc = d_q.alloc(q_cap * sizeof(TaskRef))) ||
        (rc = d_ready.alloc(q_cap * 4)) || (rc = d_ctr.alloc(4 * 8)) || (rc = d_slab.alloc((size_t)grid * slab_b)) ||
        (rc = d_clq_refs.alloc(clq_refs * sizeof(CliqueRef))) || (rc = d_clq_verts.alloc(clq_slots * 4)) ||
        (rc = d_clq_ctr.alloc(sizeof(StreamCtr))) || (rc = d_stats.alloc(sizeof(DevStats))) ||
        (rc = d_err.alloc(sizeof(int))) || (rc = d_args.alloc(sizeof(ExpandArgs))))
        return rc;

    // queue: the root at slot 0
    unsigned long long ctr[4] = {0 /*head*/, 1 /*tail*/, (unsigned long long)I.R /*pool top*/, 1 /*pending*/};
    TaskRef root{0, (uint32_t)I.R, 0, I.root_bound, 0, 0};
    unsigned int one = 1;
    QC_HIP(hipMemsetAsync(d_ready.p, 0, q_cap * 4, s));
    QC_HIP(hipMemcpyAsync(d_pool.p, I.root.data(), I.root.size() * sizeof(Vrec), hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_q.p, &root, sizeof(root), hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_ready.p, &one, 4, hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_ctr.p, ctr, sizeof(ctr), hipMemcpyHostToDevice, s));
    QC_HIP(hipMemsetAsync(d_clq_ctr.p, 0, sizeof(StreamCtr), s));
    QC_HIP(hipMemsetAsync(d_stats.p, 0, sizeof(DevStats), s));
    QC_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), s));

    ExpandArgs A;
    memset(&A, 0, sizeof(A));
    A.g = I.dg;
    A.in_verts = d_pool.as<Vrec>();
    A.small_max = (uint32_t)TL;
    A.clq_refs = d_clq_refs.as<CliqueRef>();
    A.clq_verts = d_clq_verts.as<uint32_t>();
    A.clq_ctr = d_clq_ctr.as<StreamCtr>();
    A.clq_ref_cap = clq_refs;
    A.clq_vert_cap = clq_slots;
    A.slab = d_slab.as<char>();
    A.slab_bytes = slab_b;
    A.slab_tcap = (uint32_t)I.R;
    A.stats = d_stats.as<DevStats>();
    A.error = d_err.as<int>();
    A.root_b0 = p.root_begin < 0 ? 0 : p.root_begin;
    A.root_b1 = p.root_end > 0 ? p.root_end : -1;
    A.max_depth = p.max_depth > 0 ? p.max_depth : -1;
    A.queue_mode = 1;
    A.split_chunk = p.split_chunk < 0 ? 0 : (p.split_chunk == 0 ? 16 : p.split_chunk);
    A.split_min_t = (uint32_t)TL;
    A.q_refs = d_q.as<TaskRef>();
    A.q_ready = d_ready.as<unsigned int>();
    A.q_head = d_ctr.as<unsigned long long>() + 0;
    A.q_tail = d_ctr.as<unsigned long long>() + 1;
    A.pool_top = d_ctr.as<unsigned long long>() + 2;
    A.pending = (long long*)(d_ctr.as<unsigned long long>() + 3);
    A.q_cap = q_cap;
    A.pool = d_pool.as<Vrec>();
    A.pool_cap = pool_cap;
    QC_HIP(hipMemcpyAsync(d_args.p, &A, sizeof(A), hipMemcpyHostToDevice, s));

    hipEvent_t e0, e1;
    QC_HIP(hipEventCreate(&e0));
    QC_HIP(hipEventCreate(&e1));
    QC_HIP(hipEventRecord(e0, s));
    const ExpandArgs* ap = d_args.as<ExpandArgs>();
    void* kargs[] = {(void*)&ap};
    QC_HIP(hipLaunchKernel(kern, dim3(grid), dim3(64), kargs, 0, s));
    QC_HIP(hipEventRecord(e1, s));
    QC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    QC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out.kernel_ms = ms;
    out.levels = 1;
    int err = 0;
    QC_HIP(hipMemcpy(&err, d_err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        set_error("persistent expansion failed (error flags %d: 1 capacity, 2 task too large, 4 queue hand-off)", err);
        return (err & 4) ? QC_EDEVICE : QC_EOVERFLOW;
    }
    StreamCtr cc;
    QC_HIP(hipMemcpy(&cc, d_clq_ctr.p, sizeof(cc), hipMemcpyDeviceToHost));
    std::vector<CliqueRef> hr(cc.items);
    std::vector<uint32_t> hv(cc.slots);
    if (cc.items) QC_HIP(hipMemcpy(hr.data(), d_clq_refs.p, cc.items * sizeof(CliqueRef), hipMemcpyDeviceToHost));
    if (cc.slots) QC_HIP(hipMemcpy(hv.data(), d_clq_verts.p, cc.slots * 4, hipMemcpyDeviceToHost));
    out.raw_off.reserve(cc.items + 1);
    out.raw_v.reserve(cc.slots);
    for (auto& r : hr) {
        out.raw_v.insert(out.raw_v.end(), hv.begin() + r.off, hv.begin() + r.off + r.len);
        out.raw_off.push_back(out.raw_v.size());
    }
    QC_HIP(hipMemcpy(&out.ds, d_stats.p, sizeof(DevStats), hipMemcpyDeviceToHost));
    return QC_OK;
}

}  // namespace qc
