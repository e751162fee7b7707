// driver.hip -- the plain (level-synchronous) enumeration: the structure of
// the reference's search() loop (main.cu:606-678) on the device.
//
// Tasks wait on two LIFO stacks (vertex pool in HBM, task refs mirrored on
// the host): tasks of at most `TL` vertices are expanded from LDS by
// expand_local<TL>, larger ones from per-wave HBM slabs by expand_hbm.  Each
// host iteration pops a batch from the top of one stack (large first), runs
// one launch, and pushes the children it wrote back on top, so the search
// proceeds depth-first in batches, like the reference's width-bounded levels
// with the LIFO overflow buffer (main.cu:2926-2931).  Raw cliques stream into
// an HBM buffer that is drained to the host when more than half full
// (CLIQUES_DUMP, main.cu:56).

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>

#include "qc_internal.h"

namespace qcdev {
template <int TL>
__global__ void expand_local(const ExpandArgs*);
__global__ void expand_hbm(const ExpandArgs*);
}  // namespace qcdev

namespace qc {

namespace {

struct Stack {
    DBuf verts;                    // Vrec pool
    uint64_t cap = 0;              // slots
    uint64_t top = 0;              // slots in use
    std::vector<TaskRef> refs;     // host mirror, increasing off
};

size_t slab_bytes_for(uint32_t tcap, uint32_t rw) {
    uint32_t T2 = 64;
    while (T2 < tcap) T2 <<= 1;
    size_t b = (size_t)15 * T2 * 4;
    b = (b + 15) & ~(size_t)15;
    b += (size_t)2 * rw * 8;
    return (b + 255) & ~(size_t)255;
}

}  // namespace

int enumerate_plain(const Instance& I, const qc_params& p, EnumOut& out, hipStream_t s) {
    int rc;
    out = EnumOut();
    if (I.R == 0) return QC_OK;
    const uint32_t TL = p.lds_tile == 64 ? 64 : p.lds_tile == 256 ? 256 : 128;

    // slot bound of the root's children inside the search window: branch b
    // takes root vertex R - 1 - b and writes at most min(R - b, 1 + |N2 n root|)
    // vertices
    uint64_t root_bound = 0;
    {
        const long long nb = (long long)I.root_branches;
        long long b0 = p.root_begin >= 0 ? p.root_begin : std::max(0ll, nb + (long long)p.root_begin);
        long long b1 = p.root_end > 0 ? std::min<long long>(p.root_end, nb) : nb;
        b0 = std::min(b0, nb);
        for (long long b = b0; b < b1; b++) {
            const Vrec& x = I.root[(size_t)(I.R - 1 - b)];
            root_bound += std::min<uint64_t>((uint64_t)(I.R - b), (uint64_t)I.d2r[x.id] + 1);
        }
    }

    // ---- capacities ----
    size_t freeb = 0, totalb = 0;
    QC_HIP(hipMemGetInfo(&freeb, &totalb));
    {
        int d = 0;
        QC_HIP(hipGetDevice(&d));
        freeb += pool_cached_bytes(d);
    }
    // Buffers start near what the root needs and grow on demand up to a
    // budget; task_capacity / clique_capacity pin them instead.
    const size_t budget = std::min<size_t>(freeb / 2, (size_t)96 << 30);
    const uint64_t max_slots = budget / 5 / sizeof(Vrec);
    uint64_t stg_slot_cap = p.task_capacity ? p.task_capacity
                                            : std::min<uint64_t>(max_slots, std::max<uint64_t>(8ull << 20, 2 * root_bound));
    uint64_t pool_cap = p.task_capacity ? p.task_capacity : stg_slot_cap;
    const bool grow = p.task_capacity == 0;
    uint64_t stage_refs = std::min<uint64_t>(stg_slot_cap / 2 + 4096, (uint64_t)128 << 20);
    uint64_t clq_slots = p.clique_capacity ? p.clique_capacity : (16ull << 20);
    uint64_t clq_refs = std::min<uint64_t>(clq_slots / 4 + 4096, (uint64_t)64 << 20);

    Stack st[2];
    for (int k = 0; k < 2; k++) {
        st[k].cap = pool_cap;
        if ((rc = st[k].verts.alloc(st[k].cap * sizeof(Vrec)))) return rc;
    }
    DBuf stg_refs[2], stg_verts[2];
    for (int k = 0; k < 2; k++) {
        if ((rc = stg_refs[k].alloc(stage_refs * sizeof(TaskRef)))) return rc;
        if ((rc = stg_verts[k].alloc(stg_slot_cap * sizeof(Vrec)))) return rc;
    }
    DBuf d_ctr, d_clq_ctr, d_clq_refs, d_clq_verts, d_stats, d_err, d_task_ctr, d_args, d_in_refs;
    if ((rc = d_ctr.alloc(2 * sizeof(StreamCtr)))) return rc;
    if ((rc = d_clq_ctr.alloc(sizeof(StreamCtr)))) return rc;
    if ((rc = d_clq_refs.alloc(clq_refs * sizeof(CliqueRef)))) return rc;
    if ((rc = d_clq_verts.alloc(clq_slots * 4))) return rc;
    if ((rc = d_stats.alloc(sizeof(DevStats)))) return rc;
    if ((rc = d_err.alloc(sizeof(int)))) return rc;
    if ((rc = d_task_ctr.alloc(sizeof(unsigned int)))) return rc;
    if ((rc = d_args.alloc(sizeof(ExpandArgs)))) return rc;
    const uint32_t max_batch = 1u << 16;
    if ((rc = d_in_refs.alloc((size_t)max_batch * sizeof(TaskRef)))) return rc;
    QC_HIP(hipMemsetAsync(d_clq_ctr.p, 0, sizeof(StreamCtr), s));
    QC_HIP(hipMemsetAsync(d_stats.p, 0, sizeof(DevStats), s));
    QC_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), s));

    // HBM slabs for tasks above the LDS tile (root-space addressing)
    const uint32_t slab_tcap = (uint32_t)I.R;
    const size_t slab_b = slab_bytes_for(slab_tcap, (uint32_t)I.rw);
    uint32_t n_slabs = 256;
    while (n_slabs > 1 && (size_t)n_slabs * slab_b > budget / 8) n_slabs >>= 1;
    DBuf d_slab;
    if ((rc = d_slab.alloc((size_t)n_slabs * slab_b))) return rc;

    int dev = 0;
    QC_HIP(hipGetDevice(&dev));
    int n_cu = 0;
    QC_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t small_grid_max =
        p.workgroups > 0 ? (uint32_t)p.workgroups : (uint32_t)std::max(1, n_cu) * 16;

    // ---- root ----
    // (a window's leading removals applied at once when they are many)
    RootCut cut;
    if ((rc = root_window_cut(I, p.root_begin, p.root_end, cut, s))) return rc;
    const std::vector<Vrec>& rrec = cut.cut ? cut.rec : I.root;
    const int rt = (int)rrec.size();
    {
        const int k = rt <= (int)TL ? 0 : 1;
        QC_HIP(hipMemcpyAsync(st[k].verts.p, rrec.data(), rrec.size() * sizeof(Vrec), hipMemcpyHostToDevice, s));
        TaskRef r{0, (uint32_t)rt, 0, 0, root_bound, 0, 0};
        st[k].refs.push_back(r);
        st[k].top = (uint64_t)rt;
    }

    ExpandArgs A;
    memset(&A, 0, sizeof(A));
    A.g = I.dg;
    A.static_sched = p.scheduling != 0;
    A.task_counter = d_task_ctr.as<unsigned int>();
    for (int k = 0; k < 2; k++) {
        A.out_refs[k] = stg_refs[k].as<TaskRef>();
        A.out_verts[k] = stg_verts[k].as<Vrec>();
        A.out_ref_cap[k] = stage_refs;
        A.out_vert_cap[k] = stg_slot_cap;
    }
    A.out_ctr = d_ctr.as<StreamCtr>();
    A.small_max = TL;
    A.clq_refs = d_clq_refs.as<CliqueRef>();
    A.clq_verts = d_clq_verts.as<uint32_t>();
    A.clq_ctr = d_clq_ctr.as<StreamCtr>();
    A.clq_ref_cap = clq_refs;
    A.clq_vert_cap = clq_slots;
    A.slab = d_slab.as<char>();
    A.slab_bytes = slab_b;
    A.slab_tcap = slab_tcap;
    A.slab_mwords = (uint32_t)I.rw;
    A.stats = d_stats.as<DevStats>();
    A.error = d_err.as<int>();
    A.g.root_b0 = cut.b0;
    A.g.root_b1 = cut.b1;
    A.g.root_cut = cut.cut;
    A.g.max_depth = p.max_depth > 0 ? p.max_depth : -1;

    hipEvent_t e0, e1;
    QC_HIP(hipEventCreate(&e0));
    QC_HIP(hipEventCreate(&e1));
    double kms = 0;
    std::vector<CliqueRef> hrefs;
    std::vector<uint32_t> hverts;

    auto drain = [&](bool force) -> int {
        StreamCtr cc;
        QC_HIP(hipMemcpyAsync(&cc, d_clq_ctr.p, sizeof(cc), hipMemcpyDeviceToHost, s));
        QC_HIP(hipStreamSynchronize(s));
        if (!force && cc.slots * 2 < clq_slots && cc.items * 2 < clq_refs) return QC_OK;
        if (cc.items == 0) return QC_OK;
        hrefs.resize(cc.items);
        hverts.resize(cc.slots);
        QC_HIP(hipMemcpyAsync(hrefs.data(), d_clq_refs.p, cc.items * sizeof(CliqueRef), hipMemcpyDeviceToHost, s));
        QC_HIP(hipMemcpyAsync(hverts.data(), d_clq_verts.p, cc.slots * 4, hipMemcpyDeviceToHost, s));
        QC_HIP(hipStreamSynchronize(s));
        for (auto& r : hrefs) {
            out.raw_v.insert(out.raw_v.end(), hverts.begin() + r.off, hverts.begin() + r.off + r.len);
            out.raw_off.push_back(out.raw_v.size());
        }
        QC_HIP(hipMemsetAsync(d_clq_ctr.p, 0, sizeof(StreamCtr), s));
        return QC_OK;
    };

    // QCMINE_DEBUG=1: the reference's per-level size report on stdout
    const bool debug = getenv("QCMINE_DEBUG") && atoi(getenv("QCMINE_DEBUG")) > 0;
    struct {
        uint64_t tasks = 0, buf = 0, buf_off = 0, clq = 0, clq_off = 0, child = 0, child_off = 0;
    } mx;
    std::vector<TaskRef> batch, kids;
    while (!st[0].refs.empty() || !st[1].refs.empty()) {
        const int k = st[1].refs.empty() ? 0 : 1;
        Stack& S = st[k];
        // batch from the top: bounded count, children fit the staging buffers
        const uint32_t bmax = k == 1 ? n_slabs : std::min<uint32_t>(max_batch, small_grid_max * 4);
        uint64_t vol = 0;
        size_t take = 0;
        while (take < S.refs.size() && take < bmax) {
            const TaskRef& r = S.refs[S.refs.size() - 1 - take];
            if (take > 0 && vol + r.bound > stg_slot_cap) break;
            if (r.bound > stg_slot_cap) {
                uint64_t want = std::max<uint64_t>(r.bound, 2 * stg_slot_cap);
                if (!grow || r.bound > max_slots) {
                    set_error("task of %u vertices needs %llu child slots, staging holds %llu", r.t,
                              (unsigned long long)r.bound, (unsigned long long)stg_slot_cap);
                    return QC_EOVERFLOW;
                }
                stg_slot_cap = std::min(want, max_slots);
                stage_refs = std::min<uint64_t>(stg_slot_cap / 2 + 4096, (uint64_t)128 << 20);
                for (int j = 0; j < 2; j++) {
                    stg_verts[j].release();
                    stg_refs[j].release();
                    if ((rc = stg_verts[j].alloc(stg_slot_cap * sizeof(Vrec)))) return rc;
                    if ((rc = stg_refs[j].alloc(stage_refs * sizeof(TaskRef)))) return rc;
                    A.out_refs[j] = stg_refs[j].as<TaskRef>();
                    A.out_verts[j] = stg_verts[j].as<Vrec>();
                    A.out_ref_cap[j] = stage_refs;
                    A.out_vert_cap[j] = stg_slot_cap;
                }
            }
            vol += r.bound;
            take++;
        }
        batch.assign(S.refs.end() - take, S.refs.end());
        S.refs.resize(S.refs.size() - take);
        const uint64_t base = batch.front().off;
        QC_HIP(hipMemcpyAsync(d_in_refs.p, batch.data(), take * sizeof(TaskRef), hipMemcpyHostToDevice, s));
        A.in_refs = d_in_refs.as<TaskRef>();
        A.in_verts = S.verts.as<Vrec>();
        A.n_in = (uint32_t)take;
        QC_HIP(hipMemcpyAsync(d_args.p, &A, sizeof(A), hipMemcpyHostToDevice, s));
        QC_HIP(hipMemsetAsync(d_ctr.p, 0, 2 * sizeof(StreamCtr), s));
        QC_HIP(hipMemsetAsync(d_task_ctr.p, 0, sizeof(unsigned int), s));
        QC_HIP(hipEventRecord(e0, s));
        if (k == 1) {
            uint32_t grid = std::min<uint32_t>((uint32_t)take, n_slabs);
            hipLaunchKernelGGL(qcdev::expand_hbm, dim3(grid), dim3(64), 0, s, d_args.as<ExpandArgs>());
        } else {
            uint32_t grid = std::min<uint32_t>((uint32_t)take, small_grid_max);
            if (TL == 64)
                hipLaunchKernelGGL(qcdev::expand_local<64>, dim3(grid), dim3(64), 0, s, d_args.as<ExpandArgs>());
            else if (TL == 128)
                hipLaunchKernelGGL(qcdev::expand_local<128>, dim3(grid), dim3(64), 0, s, d_args.as<ExpandArgs>());
            else
                hipLaunchKernelGGL(qcdev::expand_local<256>, dim3(grid), dim3(64), 0, s, d_args.as<ExpandArgs>());
        }
        QC_HIP(hipGetLastError());
        QC_HIP(hipEventRecord(e1, s));
        StreamCtr oc[2];
        int err = 0;
        QC_HIP(hipMemcpyAsync(oc, d_ctr.p, sizeof(oc), hipMemcpyDeviceToHost, s));
        QC_HIP(hipMemcpyAsync(&err, d_err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        QC_HIP(hipStreamSynchronize(s));
        float ms = 0;
        QC_HIP(hipEventElapsedTime(&ms, e0, e1));
        kms += ms;
        out.levels++;
        if (err) {
            set_error("expansion buffers overflowed (error flags %d); raise task_capacity / clique_capacity", err);
            return QC_EOVERFLOW;
        }
        S.top = base;
        // push children: staging stream j -> stack j
        for (int j = 0; j < 2; j++) {
            if (oc[j].items == 0) continue;
            Stack& D = st[j];
            if (D.top + oc[j].slots > D.cap && grow && D.cap < max_slots) {
                uint64_t ncap = std::min<uint64_t>(max_slots, std::max<uint64_t>(2 * D.cap, D.top + oc[j].slots));
                DBuf nb;
                if ((rc = nb.alloc(ncap * sizeof(Vrec)))) return rc;
                if (D.top) QC_HIP(hipMemcpyAsync(nb.p, D.verts.p, D.top * sizeof(Vrec), hipMemcpyDeviceToDevice, s));
                QC_HIP(hipStreamSynchronize(s));
                D.verts.swap(nb);
                D.cap = ncap;
            }
            if (D.top + oc[j].slots > D.cap) {
                set_error("task stack overflow (%llu + %llu slots > %llu)", (unsigned long long)D.top,
                          (unsigned long long)oc[j].slots, (unsigned long long)D.cap);
                return QC_EOVERFLOW;
            }
            kids.resize(oc[j].items);
            QC_HIP(hipMemcpyAsync(kids.data(), stg_refs[j].p, oc[j].items * sizeof(TaskRef), hipMemcpyDeviceToHost, s));
            QC_HIP(hipMemcpyAsync(D.verts.as<Vrec>() + D.top, stg_verts[j].p, oc[j].slots * sizeof(Vrec),
                                  hipMemcpyDeviceToDevice, s));
            QC_HIP(hipStreamSynchronize(s));
            std::sort(kids.begin(), kids.end(), [](const TaskRef& x, const TaskRef& y) { return x.off < y.off; });
            for (auto& r : kids) {
                r.off += D.top;
                D.refs.push_back(r);
            }
            D.top += oc[j].slots;
        }
        if (debug) {
            // the reference's debug size report (main.cu:2635-2646): level,
            // tasks expanded (records, vertex slots), children written,
            // records waiting on the host stacks, raw cliques held
            StreamCtr cc;
            QC_HIP(hipMemcpyAsync(&cc, d_clq_ctr.p, sizeof(cc), hipMemcpyDeviceToHost, s));
            QC_HIP(hipStreamSynchronize(s));
            uint64_t bslots = 0;
            for (const TaskRef& r : batch) bslots += r.t;
            if (out.levels == 1) bslots += (uint64_t)cut.cut;  // the root as built
            const uint64_t wrefs = st[0].refs.size() + st[1].refs.size(), wslots = st[0].top + st[1].top;
            printf("L: %llu T1: %llu %llu T2: %llu %llu B: %llu %llu C: %llu %llu\n\n", (unsigned long long)out.levels,
                   (unsigned long long)take, (unsigned long long)bslots,
                   (unsigned long long)(oc[0].items + oc[1].items), (unsigned long long)(oc[0].slots + oc[1].slots),
                   (unsigned long long)wrefs, (unsigned long long)wslots, cc.items, cc.slots);
            mx.tasks = std::max<uint64_t>(mx.tasks, bslots);
            mx.buf = std::max<uint64_t>(mx.buf, wslots);
            mx.buf_off = std::max<uint64_t>(mx.buf_off, wrefs);
            mx.clq = std::max<uint64_t>(mx.clq, cc.slots);
            mx.clq_off = std::max<uint64_t>(mx.clq_off, cc.items);
            mx.child = std::max<uint64_t>(mx.child, oc[0].slots + oc[1].slots);
            mx.child_off = std::max<uint64_t>(mx.child_off, oc[0].items + oc[1].items);
        }
        if ((rc = drain(false))) return rc;
    }
    if ((rc = drain(true))) return rc;
    if (debug) {
        // print_maxes (main.cu:2893-2907): the largest of each buffer over the run
        printf("\nTASKS SIZE: %llu\nBUFFER SIZE: %llu\nBUFFER OFFSET SIZE: %llu\nCLIQUES SIZE: %llu\n"
               "CLIQUES OFFSET SIZE: %llu\nWTASKS SIZE: %llu\nWTASKS OFFSET SIZE: %llu\n\n",
               (unsigned long long)mx.tasks, (unsigned long long)mx.buf, (unsigned long long)mx.buf_off,
               (unsigned long long)mx.clq, (unsigned long long)mx.clq_off, (unsigned long long)mx.child,
               (unsigned long long)mx.child_off);
        fflush(stdout);
    }
    QC_HIP(hipMemcpyAsync(&out.ds, d_stats.p, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    root_cut_counters(cut, I.R, TL, out.ds);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out.kernel_ms = kms;
    return QC_OK;
}

}  // namespace qc
