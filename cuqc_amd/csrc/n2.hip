// n2.hip -- the serialized format's N2 lists built on the GPU and written to
// the file in chunks (qc_graph_write_serialized_device).
//
// adjToSer.cpp:115-156 writes, for every vertex v, N2(v) = N(v) u N(N(v))
// minus v in ascending order, and the serialized header (adjToSer.cpp:
// 160-200, read back by main.cu:96-122) names their total L2 before the first
// list.  On graphs of a million vertices the lists hold billions of entries,
// more than the host's flag-array loop can build in reasonable time or keep
// in memory.  Here:
//   1. count pass: one workgroup per vertex marks N(v) and the lists of its
//      neighbours in a visited bitmap over all n vertices held in LDS, clears
//      v, and popcounts the map: |N2(v)|.  The host's prefix sums give the
//      N2 offsets and L2;
//   2. write passes, over chunks of consecutive vertices whose lists fit the
//      output buffer: the same map, read out in ascending order (each thread
//      owns a run of map words; a block scan of their popcounts places every
//      thread's ids), copied to the host and formatted into the file.
// The host holds one chunk of entries at a time.  The map is n bits of LDS,
// so n is limited by the LDS a workgroup may hold (160 KB: 1,310,720
// vertices less the kernel's static LDS).

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "qc_internal.h"

namespace qcdev {

constexpr int N2_THREADS = 256;

template <bool WRITE>
__global__ __launch_bounds__(N2_THREADS) void n2_pass(const uint64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                      int32_t v0, int32_t v1, uint32_t map_words,
                                                      uint32_t* __restrict__ cnt, const uint64_t* __restrict__ off2,
                                                      uint64_t base, uint32_t* __restrict__ out) {
    extern __shared__ unsigned long long map[];
    __shared__ uint32_t s_part[N2_THREADS / 64];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wv = tid >> 6, nwv = nt >> 6;
    // each thread reads out a run of consecutive map words (ascending ids)
    const uint32_t per = (map_words + (uint32_t)nt - 1) / (uint32_t)nt;
    const uint32_t w0 = min(map_words, (uint32_t)tid * per), w1 = min(map_words, w0 + per);
    for (int32_t v = v0 + (int32_t)blockIdx.x; v < v1; v += (int32_t)gridDim.x) {
        for (uint32_t w = tid; w < map_words; w += nt) map[w] = 0;
        __syncthreads();
        const uint64_t b0 = off[v], b1 = off[v + 1];
        for (uint64_t j = b0 + tid; j < b1; j += nt) {
            const uint32_t u = adj[j];
            atomicOr(&map[u >> 6], 1ull << (u & 63));
        }
        // N(u) for the neighbours u: wave wv walks neighbours wv, wv + nwv, ...
        for (uint64_t j = b0 + wv; j < b1; j += nwv) {
            const uint32_t u = adj[j];
            const uint64_t e0 = off[u], e1 = off[u + 1];
            for (uint64_t k = e0 + lane; k < e1; k += 64) {
                const uint32_t x = adj[k];
                atomicOr(&map[x >> 6], 1ull << (x & 63));
            }
        }
        __syncthreads();
        if (tid == 0) map[(uint32_t)v >> 6] &= ~(1ull << ((uint32_t)v & 63));
        __syncthreads();
        uint32_t c = 0;
        for (uint32_t w = w0; w < w1; w++) c += (uint32_t)__popcll(map[w]);
        // block exclusive scan of the per-thread counts
        uint32_t incl = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) s_part[wv] = incl;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int q = 0; q < nwv; q++) {
            if (q < wv) before += s_part[q];
            total += s_part[q];
        }
        if (!WRITE) {
            if (tid == 0) cnt[v] = total;
        } else {
            uint64_t o = off2[v] - base + before + incl - c;
            for (uint32_t w = w0; w < w1; w++) {
                unsigned long long x = map[w];
                while (x) {
                    out[o++] = w * 64u + (uint32_t)__builtin_ctzll(x);
                    x &= x - 1ull;
                }
            }
        }
        __syncthreads();  // s_part and the map are reused by the next vertex
    }
}

}  // namespace qcdev

namespace qc {

namespace {

// text output: "%u"/"%llu" numbers, one space between, newline per record
struct TextOut {
    FILE* f;
    std::string buf;
    bool ok = true;
    explicit TextOut(FILE* fp) : f(fp) { buf.reserve(1 << 24); }
    void num(uint64_t x) {
        char t[24];
        int k = 0;
        do {
            t[k++] = (char)('0' + x % 10);
            x /= 10;
        } while (x);
        while (k) buf.push_back(t[--k]);
        if (buf.size() >= (1u << 24) - 32) flush();
    }
    void ch(char c) { buf.push_back(c); }
    void flush() {
        if (!buf.empty() && fwrite(buf.data(), 1, buf.size(), f) != buf.size()) ok = false;
        buf.clear();
    }
    template <class T>
    void record(const T* a, size_t k) {
        for (size_t i = 0; i < k; i++) {
            if (i) ch(' ');
            num((uint64_t)a[i]);
        }
        ch('\n');
    }
};

#define QC_N2(x)                                                 \
    do {                                                         \
        if ((x) != hipSuccess) {                                 \
            set_error("device N2 build: %s failed", #x);         \
            return QC_EDEVICE;                                   \
        }                                                        \
    } while (0)

int write_n2_file(const qc_graph* g, const char* path, hipStream_t s) {
    const int32_t n = g->n;
    const uint32_t map_words = (uint32_t)(((uint64_t)n + 63) / 64);
    const size_t lds_bytes = std::max<size_t>((size_t)map_words * 8, 8);
    int dev = 0, lds_max = 0, n_cu = 0;
    QC_N2(hipGetDevice(&dev));
    QC_N2(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    QC_N2(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t static_lds = sizeof(uint32_t) * (qcdev::N2_THREADS / 64);
    if (lds_bytes + static_lds > (size_t)lds_max) {
        set_error("device N2 build: a visited map of %d vertices needs %zu bytes of LDS, the device gives a "
                  "workgroup %d",
                  n, lds_bytes, lds_max);
        return QC_EINVAL;
    }
    if (lds_bytes > 48 * 1024) {
        QC_N2(hipFuncSetAttribute((const void*)qcdev::n2_pass<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_bytes));
        QC_N2(hipFuncSetAttribute((const void*)qcdev::n2_pass<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_bytes));
    }
    const uint64_t* d_off = nullptr;
    const uint32_t* d_adj = nullptr;
    int rc;
    if ((rc = graph_device_arrays(g, &d_off, &d_adj, s))) return rc;

    // ---- count pass ----
    std::vector<uint64_t> off2((size_t)n + 1, 0);
    DBuf d_cnt, d_off2, d_out;
    if ((rc = d_cnt.alloc((size_t)std::max(n, 1) * 4, s)) || (rc = d_off2.alloc(((size_t)n + 1) * 8, s))) return rc;
    // enough workgroups to fill every CU several times over; each loops over
    // vertices v = blockIdx, blockIdx + gridDim, ...
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)n_cu * 16));
    if (n > 0) {
        hipLaunchKernelGGL(qcdev::n2_pass<false>, dim3(blocks), dim3(qcdev::N2_THREADS), lds_bytes, s, d_off, d_adj,
                           0, n, map_words, d_cnt.as<uint32_t>(), (const uint64_t*)nullptr, (uint64_t)0,
                           (uint32_t*)nullptr);
        QC_N2(hipGetLastError());
        std::vector<uint32_t> cnt(n);
        QC_N2(hipMemcpyAsync(cnt.data(), d_cnt.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        QC_N2(hipStreamSynchronize(s));
        for (int32_t v = 0; v < n; v++) off2[v + 1] = off2[v] + cnt[v];
        QC_N2(hipMemcpyAsync(d_off2.p, off2.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, s));
    }
    const uint64_t L2 = off2[n];
    const uint64_t E = g->adj.size();

    FILE* f = fopen(path, "w");
    if (!f) {
        set_error("cannot open %s for write", path);
        return QC_EIO;
    }
    TextOut t(f);
    t.num((uint64_t)(int64_t)n);  // n >= 0
    t.ch('\n');
    t.num(E);
    t.ch('\n');
    t.num(L2);
    t.ch('\n');
    t.record(g->adj.data(), (size_t)E);
    t.record(g->off.data(), (size_t)n + 1);

    // ---- write passes: chunks of vertices whose lists fit the buffer ----
    // entries per chunk: 2^26 (256 MB), or QCMINE_N2_CHUNK; at least n, so
    // every vertex's list fits one chunk
    uint64_t cap = (uint64_t)1 << 26;
    if (const char* e = getenv("QCMINE_N2_CHUNK")) {
        const long long v = atoll(e);
        if (v > 0) cap = (uint64_t)v;
    }
    cap = std::max<uint64_t>(cap, (uint64_t)std::max(n, 1));
    if (L2 > 0 && (rc = d_out.alloc((size_t)std::min(cap, L2) * 4, s))) {
        fclose(f);
        return rc;
    }
    std::vector<uint32_t> host;
    bool first = true;
    for (int32_t v0 = 0; v0 < n;) {
        int32_t v1 = v0;
        while (v1 < n && off2[v1 + 1] - off2[v0] <= cap) v1++;
        const uint64_t k = off2[v1] - off2[v0];
        if (k > 0) {
            const int cb = (int)std::max<int64_t>(1, std::min<int64_t>(v1 - v0, (int64_t)n_cu * 16));
            hipLaunchKernelGGL(qcdev::n2_pass<true>, dim3(cb), dim3(qcdev::N2_THREADS), lds_bytes, s, d_off, d_adj,
                               v0, v1, map_words, (uint32_t*)nullptr, d_off2.as<uint64_t>(), off2[v0],
                               d_out.as<uint32_t>());
            if (hipGetLastError() != hipSuccess) {
                fclose(f);
                set_error("device N2 build: the write pass failed to launch");
                return QC_EDEVICE;
            }
            host.resize(k);
            if (hipMemcpyAsync(host.data(), d_out.p, k * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                fclose(f);
                set_error("device N2 build: the write pass failed");
                return QC_EDEVICE;
            }
            for (uint64_t i = 0; i < k; i++) {
                if (!first) t.ch(' ');
                first = false;
                t.num(host[i]);
            }
        }
        v0 = v1;
    }
    t.ch('\n');
    t.record(off2.data(), (size_t)n + 1);
    t.flush();
    if (fclose(f) != 0 || !t.ok) {
        set_error("write to %s failed", path);
        return QC_EIO;
    }
    return QC_OK;
}

}  // namespace

int write_serialized_device(const qc_graph* g, const char* path) {
    hipStream_t s;
    if (hipStreamCreate(&s) != hipSuccess) {
        set_error("device N2 build: no HIP stream");
        return QC_EDEVICE;
    }
    int rc = write_n2_file(g, path, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return rc;
}

}  // namespace qc
