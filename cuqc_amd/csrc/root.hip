// root.hip -- root task construction (initialize_tasks, main.cu:846-998) and
// the device tables the expansion kernels read.
//
// Host part: the degree / |N2| filter (main.cu:872-885), the k-core peel
// (main.cu:889-949), the cover vertex and covered set (main.cu:953-975) and
// the Q-ordered root record (main.cu:978-995).
// Device part: for every root vertex, its rows of the root-space 1-hop and
// 2-hop bit matrices and its full-graph |N2| (the third Q-order key), with
// the 2-hop relation of adjToSer.cpp:115-156: N(v) u N(N(v)) minus v.
//
// Root-local ids are assigned in increasing input-id order, so comparing
// root-local ids orders vertices exactly like comparing input ids (the last
// Q-order key, main.cu:2098-2101).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "qc_internal.h"

namespace qcdev {

// One wave per root vertex.  vis: per-wave visited bitmap over all n input
// vertices (zero on entry and on exit).
__global__ __launch_bounds__(64) void root_rows(const uint64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                const uint64_t* __restrict__ off2, const uint32_t* __restrict__ adj2,
                                                const int32_t* __restrict__ loc, const uint32_t* __restrict__ orig,
                                                int32_t R, int32_t rw, uint64_t* A1, uint64_t* A2, int32_t* n2out,
                                                unsigned long long* vis_all, uint64_t vis_words, int has_n2) {
    const int lane = threadIdx.x & 63;
    unsigned long long* vis = vis_all + (size_t)blockIdx.x * vis_words;
    for (int a = blockIdx.x; a < R; a += gridDim.x) {
        const uint32_t v = orig[a];
        unsigned long long* row1 = (unsigned long long*)(A1 + (size_t)a * rw);
        unsigned long long* row2 = (unsigned long long*)(A2 + (size_t)a * rw);
        const uint64_t b0 = off[v], b1 = off[v + 1];
        for (uint64_t j = b0 + lane; j < b1; j += 64) {
            int la = loc[adj[j]];
            if (la >= 0) atomicOr(&row1[la >> 6], 1ull << (la & 63));
        }
        if (has_n2) {
            for (uint64_t j = off2[v] + lane; j < off2[v + 1]; j += 64) {
                int la = loc[adj2[j]];
                if (la >= 0) atomicOr(&row2[la >> 6], 1ull << (la & 63));
            }
            if (lane == 0) n2out[a] = (int32_t)(off2[v + 1] - off2[v]);
            __syncthreads();
            continue;
        }
        if (lane == 0) atomicOr(&vis[v >> 6], 1ull << (v & 63));
        __syncthreads();
        int cnt = 0;
        for (uint64_t j = b0 + lane; j < b1; j += 64) {
            uint32_t u = adj[j];
            unsigned long long bit = 1ull << (u & 63);
            unsigned long long old = atomicOr(&vis[u >> 6], bit);
            if (!(old & bit)) {
                cnt++;
                int la = loc[u];
                if (la >= 0) atomicOr(&row2[la >> 6], 1ull << (la & 63));
            }
        }
        for (uint64_t j = b0; j < b1; j++) {
            uint32_t u = adj[j];
            for (uint64_t k = off[u] + lane; k < off[u + 1]; k += 64) {
                uint32_t x = adj[k];
                unsigned long long bit = 1ull << (x & 63);
                unsigned long long old = atomicOr(&vis[x >> 6], bit);
                if (!(old & bit)) {
                    cnt++;
                    int la = loc[x];
                    if (la >= 0) atomicOr(&row2[la >> 6], 1ull << (la & 63));
                }
            }
        }
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
        if (lane == 0) n2out[a] = cnt;
        __syncthreads();
        // restore the visited bitmap to zero
        if (lane == 0) vis[v >> 6] = 0;
        for (uint64_t j = b0 + lane; j < b1; j += 64) vis[adj[j] >> 6] = 0;
        for (uint64_t j = b0; j < b1; j++) {
            uint32_t u = adj[j];
            for (uint64_t k = off[u] + lane; k < off[u + 1]; k += 64) vis[adj[k] >> 6] = 0;
        }
        __syncthreads();
    }
}

// The same rows and counts with the visited bitmap and both rows in LDS
// (dynamic: ceil(n/64) + 2 rw words), for graphs whose bitmap fits.  Rows are
// written whole, so A1 / A2 need no clearing.  NT threads share one vertex:
// 256, or 1024 when the map leaves room for one workgroup per CU (then the
// workgroup's waves are all the CU holds, and its loads of neighbour lists
// are what hides the memory latency).
constexpr int ROWS_THREADS = 256;
constexpr int ROWS_THREADS_WIDE = 1024;
constexpr int ROWS_BIG_MAX = 512;  // long neighbour lists listed per vertex (LDS)

__device__ __forceinline__ uint64_t rdl64(uint64_t x, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

template <bool BYTES, int NT>
__global__ __launch_bounds__(NT) void root_rows_lds(const uint64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                    const int32_t* __restrict__ loc, const uint32_t* __restrict__ orig,
                                                    int32_t R, int32_t rw, int32_t n, uint64_t* A1, uint64_t* A2,
                                                    int32_t* n2out) {
    // vis: {v} u N(v) u N(N(v)) over all n vertices, as a byte map (BYTES:
    // plain stores of 1, no read-modify-write) or a bitmap (atomics without
    // return); r1: root-space row of N(v).  |N2(v)| and the A2 row are read
    // off the map afterwards.
    // NT / 64 waves share one vertex: each takes every nwv-th group
    // of neighbour lists and every nwv-th block of row words.  Lists of NT
    // entries or more (the hubs of a power-law graph) are walked by all the
    // workgroup's threads together instead: in a group they kept one wave
    // busy while the others had finished (on config 4 the busiest wave of a
    // vertex held 5.5x the average share).
    extern __shared__ unsigned long long sh[];
    __shared__ int s_cnt;
    __shared__ int s_nbig;
    __shared__ uint32_t s_big[ROWS_BIG_MAX];
    const int lane = threadIdx.x & 63;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int nw = BYTES ? (n + 7) / 8 : (n + 63) / 64;  // 64-bit words of the map
    unsigned long long* vis = sh;
    unsigned char* vis8 = (unsigned char*)sh;
    unsigned long long* r1 = sh + nw;
    auto mark = [&](uint32_t x) {
        if (BYTES)
            vis8[x] = 1;
        else
            atomicOr(&vis[x >> 6], 1ull << (x & 63));
    };
    auto seen = [&](uint32_t x) -> bool { return BYTES ? vis8[x] != 0 : ((vis[x >> 6] >> (x & 63)) & 1ull); };
    constexpr int G = 8;  // neighbours u whose lists are walked together
    for (int a = blockIdx.x; a < R; a += gridDim.x) {
        const uint32_t v = orig[a];
        for (int w = tid; w < nw; w += nt) vis[w] = 0;
        for (int w = tid; w < rw; w += nt) r1[w] = 0;
        if (tid == 0) {
            s_cnt = 0;
            s_nbig = 0;
        }
        __syncthreads();
        if (tid == 0) mark(v);
        const uint64_t b0 = off[v], b1 = off[v + 1];
        for (uint64_t j = b0 + tid; j < b1; j += nt) {
            const uint32_t u = adj[j];
            mark(u);
            const int la = loc[u];
            if (la >= 0) atomicOr(&r1[la >> 6], 1ull << (la & 63));
            if (off[u + 1] - off[u] >= (uint64_t)NT) {
                const int q = atomicAdd(&s_nbig, 1);
                if (q < ROWS_BIG_MAX) s_big[q] = u;
            }
        }
        __syncthreads();
        // the long lists, one after another, every thread on each; when more
        // than ROWS_BIG_MAX were found the groups below walk all long lists
        // again (marking twice changes nothing)
        const int nbig = min(s_nbig, ROWS_BIG_MAX);
        const bool big_done = s_nbig <= ROWS_BIG_MAX;
        for (int q = 0; q < nbig; q++) {
            const uint32_t u = s_big[q];
            const uint64_t e1 = off[u + 1];
            for (uint64_t k = off[u] + tid; k < e1; k += nt) mark(adj[k]);
        }
        // N(u) for G neighbours u at a time as one flat range of entries:
        // lane i < G holds u_i's list start and the prefix of list lengths
        for (uint64_t j0 = b0 + (uint64_t)wv * G; j0 < b1; j0 += (uint64_t)nwv * G) {
            uint64_t st = 0, len = 0;
            if (lane < G && j0 + lane < b1) {
                const uint32_t u = adj[j0 + lane];
                st = off[u];
                len = off[u + 1] - st;
                if (big_done && len >= (uint64_t)NT) len = 0;
            }
            uint64_t pre = len;  // inclusive prefix over lanes 0..G-1
            for (int o = 1; o < G; o <<= 1) {
                const uint64_t y = __shfl_up(pre, o, 64);
                if (lane >= o) pre += y;
            }
            // segment ends and starts as wave-uniform values (the G lists
            // of one group hold far fewer than 2^32 entries)
            uint32_t e[G];
            uint64_t sv[G];
#pragma unroll
            for (int q = 0; q < G; q++) {
                e[q] = (uint32_t)rdl64(pre, q);
                sv[q] = rdl64(st, q);
            }
            const uint32_t total = e[G - 1];
#pragma unroll 8
            for (uint32_t p = lane; p < total; p += 64) {
                uint64_t s0 = sv[0];
                uint32_t base = 0;
#pragma unroll
                for (int q = 0; q < G - 1; q++)
                    if (p >= e[q]) {
                        s0 = sv[q + 1];
                        base = e[q];
                    }
                mark(adj[s0 + (p - base)]);
            }
        }
        __syncthreads();
        int cnt = 0;
        for (int w = tid; w < nw; w += nt) cnt += __popcll(vis[w]);  // bytes are 0 or 1
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
        if (lane == 0) atomicAdd(&s_cnt, cnt);
        for (int w = tid; w < rw; w += nt) A1[(size_t)a * rw + w] = r1[w];
        // A2 row: bit b iff orig[b] was reached (b != a), one ballot per
        // word; eight words per step keep eight orig loads in flight
        for (int w0 = wv * 8; w0 < rw; w0 += nwv * 8) {
            uint32_t xs[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int b = (w0 + k) * 64 + lane;
                xs[k] = (w0 + k < rw && b < R) ? orig[b] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int b = (w0 + k) * 64 + lane;
                const unsigned long long word = __ballot(xs[k] != 0xFFFFFFFFu && b != a && seen(xs[k]));
                if (lane == 0 && w0 + k < rw) A2[(size_t)a * rw + w0 + k] = word;
            }
        }
        __syncthreads();
        if (tid == 0) n2out[a] = s_cnt - 1;  // v itself is in the map
    }
}

// kept root vertices against the removed ones: one wave per kept vertex,
// cnt[i] = |N(ids[i]) n removed| from its A1 row and the removal mask
__global__ __launch_bounds__(64) void root_cut_deg(const uint64_t* __restrict__ A1, int32_t rw,
                                                   const uint32_t* __restrict__ ids, int32_t nk,
                                                   const uint64_t* __restrict__ mask, int32_t* __restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    for (int i = blockIdx.x; i < nk; i += gridDim.x) {
        const uint64_t* row = A1 + (size_t)ids[i] * rw;
        int c = 0;
        for (int w = lane; w < rw; w += 64) c += __popcll(row[w] & mask[w]);
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) cnt[i] = c;
    }
}

// row popcounts: |N2(a) n root| and |N(a) n root|
__global__ __launch_bounds__(64) void root_d2(const uint64_t* __restrict__ A2, const uint64_t* __restrict__ A1,
                                              int32_t R, int32_t rw, int32_t* d2r, int32_t* d1r) {
    const int lane = threadIdx.x & 63;
    for (int a = blockIdx.x; a < R; a += gridDim.x) {
        int s = 0, s1 = 0;
        for (int w = lane; w < rw; w += 64) {
            s += __popcll(A2[(size_t)a * rw + w]);
            s1 += __popcll(A1[(size_t)a * rw + w]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            s += __shfl_xor(s, o, 64);
            s1 += __shfl_xor(s1, o, 64);
        }
        if (lane == 0) {
            d2r[a] = s;
            d1r[a] = s1;
        }
    }
}

}  // namespace qcdev

namespace qc {

namespace {

// min(|N2(v)|, cap) over the full graph: the walk stops at the cap-th
// distinct vertex (a neighbour that is a hub would otherwise cost its whole
// list for a vertex that passed long before)
long n2_capped(const qc_graph* g, int v, std::vector<int>& mark, long cap) {
    if (g->has_n2) return std::min<long>((long)(g->off2[v + 1] - g->off2[v]), cap);
    long cnt = 0;
    mark[v] = v;
    for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
        uint32_t u = g->adj[j];
        if (mark[u] != v) {
            mark[u] = v;
            if (++cnt >= cap) return cnt;
        }
        for (uint64_t k = g->off[u]; k < g->off[u + 1]; k++) {
            uint32_t x = g->adj[k];
            if (mark[x] != v) {
                mark[x] = v;
                if (++cnt >= cap) return cnt;
            }
        }
    }
    return cnt;
}

inline int label_rank(int l) { return l == 1 ? 0 : l == 2 ? 1 : l == 0 ? 2 : l == 3 ? 3 : 4; }

}  // namespace

namespace {
struct DevCSR {
    DBuf off, adj, off2, adj2;
};
void free_devcsr(void* p) { delete (DevCSR*)p; }

// the graph's arrays in HBM, uploaded once per graph (the batch builders
// run on several host threads and a batch may hold one graph twice: the
// check and the upload happen under one lock)
int device_csr(const qc_graph* g, DevCSR** out, hipStream_t s) {
    static std::mutex upload_mu;
    std::lock_guard<std::mutex> lock(upload_mu);
    qc_graph* mg = const_cast<qc_graph*>(g);
    if (!mg->dev) {
        DevCSR* d = new DevCSR;
        int rc;
        if ((rc = d->off.alloc(g->off.size() * 8)) || (rc = d->adj.alloc(g->adj.size() * 4))) {
            delete d;
            return rc;
        }
        if (hipMemcpyAsync(d->off.p, g->off.data(), g->off.size() * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(d->adj.p, g->adj.data(), g->adj.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess) {
            delete d;
            set_error("graph upload failed");
            return QC_EDEVICE;
        }
        if (g->has_n2) {
            if ((rc = d->off2.alloc(g->off2.size() * 8)) || (rc = d->adj2.alloc(g->adj2.size() * 4))) {
                delete d;
                return rc;
            }
            if (hipMemcpyAsync(d->off2.p, g->off2.data(), g->off2.size() * 8, hipMemcpyHostToDevice, s) !=
                    hipSuccess ||
                hipMemcpyAsync(d->adj2.p, g->adj2.data(), g->adj2.size() * 4, hipMemcpyHostToDevice, s) !=
                    hipSuccess) {
                delete d;
                set_error("graph upload failed");
                return QC_EDEVICE;
            }
        }
        if (hipStreamSynchronize(s) != hipSuccess) {
            delete d;
            set_error("graph upload failed");
            return QC_EDEVICE;
        }
        mg->dev = d;
        mg->dev_free = free_devcsr;
    }
    *out = (DevCSR*)mg->dev;
    return QC_OK;
}
}  // namespace

int upload_graph(const qc_graph* g, hipStream_t s) {
    DevCSR* d = nullptr;
    return device_csr(g, &d, s);
}

int graph_device_arrays(const qc_graph* g, const uint64_t** off, const uint32_t** adj, hipStream_t s) {
    DevCSR* d = nullptr;
    const int rc = device_csr(g, &d, s);
    if (rc) return rc;
    *off = d->off.as<uint64_t>();
    *adj = d->adj.as<uint32_t>();
    return QC_OK;
}

namespace {
int build_instance_impl(const qc_graph* g, double gamma, int ms, Instance& I, hipStream_t s) {
    const auto t_start = std::chrono::steady_clock::now();
    const int n = g->n;
    I.n = n;
    I.gamma = gamma;
    I.min_size = ms;
    build_tables(gamma, n, ms, I.mindeg, I.ubq);
    const int thr = I.mindeg[ms];

    // ---- filter (main.cu:872-885) ----
    std::vector<int32_t> lab(n), exd(n);
    std::vector<int32_t> rem;
    std::vector<int> mark(n, -1);
    for (int i = 0; i < n; i++) {
        int deg = (int)(g->off[i + 1] - g->off[i]);
        exd[i] = deg;
        bool pass = deg >= thr;
        if (pass && deg < ms - 1) pass = n2_capped(g, i, mark, ms - 1) >= ms - 1;
        if (pass && g->has_n2) pass = (long)(g->off2[i + 1] - g->off2[i]) >= ms - 1;
        lab[i] = pass ? 0 : -1;
        if (pass) rem.push_back(i);
    }
    const auto t_filter = std::chrono::steady_clock::now();
    // ---- peel (main.cu:889-949) ----
    // The reference's rounds and queue compute the core in which every
    // vertex has at least thr neighbours inside it (unique, whatever the
    // removal order), with exd = the degree inside it.  Here: degrees inside
    // the filtered set (host threads over the vertices), then one queue that
    // only walks the lists of removed filtered vertices.
    {
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const size_t nrem = rem.size();
        const unsigned nth = nrem > 65536 ? hw : 1u;
        auto count = [&](size_t lo, size_t hi) {
            for (size_t k = lo; k < hi; k++) {
                const int v = rem[k];
                int d = 0;
                for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) d += lab[g->adj[j]] == 0;
                exd[v] = d;
            }
        };
        if (nth == 1) {
            count(0, nrem);
        } else {
            std::vector<std::thread> th;
            for (unsigned t = 0; t < nth; t++)
                th.emplace_back(count, nrem * t / nth, nrem * (t + 1) / nth);
            for (auto& x : th) x.join();
        }
    }
    const auto t_count = std::chrono::steady_clock::now();
    std::vector<int32_t> queue;
    for (int v : rem)
        if (exd[v] < thr) {
            lab[v] = -1;
            queue.push_back(v);
        }
    for (size_t k = 0; k < queue.size(); k++) {
        const int v = queue[k];
        for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
            const uint32_t u = g->adj[j];
            if (lab[u] == 0 && --exd[u] < thr) {
                lab[u] = -1;
                queue.push_back(u);
            }
        }
    }
    const auto t_host = std::chrono::steady_clock::now();
    // ---- root-local ids ----
    I.orig.clear();
    std::vector<int32_t> loc(n, -1);
    for (int i = 0; i < n; i++)
        if (lab[i] == 0) {
            loc[i] = (int32_t)I.orig.size();
            I.orig.push_back((uint32_t)i);
        }
    const int R = (int)I.orig.size();
    I.R = R;
    I.rw = (R + 63) / 64;
    I.root.clear();
    if (R == 0) return QC_OK;

    // ---- device tables ----
    int rc;
    const size_t mat_bytes = (size_t)R * I.rw * 8;
    if ((rc = I.A1.alloc(mat_bytes))) return rc;
    if ((rc = I.A2.alloc(mat_bytes))) return rc;
    // LDS form: visited map + the 1-hop row of one root vertex per wave; a
    // byte map (plain stores) up to 32K vertices, a bitmap (atomics) above
    const bool byte_map = n <= 32768;
    const size_t map_words = byte_map ? ((size_t)n + 7) / 8 : ((size_t)n + 63) / 64;
    const size_t lds_bytes = (map_words + (size_t)I.rw) * 8;
    int lds_max = 0, cur_dev = 0;
    QC_HIP(hipGetDevice(&cur_dev));
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, cur_dev) != hipSuccess)
        lds_max = 64 * 1024;
    // one workgroup per CU when two maps do not fit: then 16 waves per vertex
    const bool rows_wide = !byte_map && lds_bytes * 2 > (size_t)lds_max;
    const int rows_threads = rows_wide ? qcdev::ROWS_THREADS_WIDE : qcdev::ROWS_THREADS;
    const void* rows_kernel = byte_map    ? (const void*)qcdev::root_rows_lds<true, qcdev::ROWS_THREADS>
                              : rows_wide ? (const void*)qcdev::root_rows_lds<false, qcdev::ROWS_THREADS_WIDE>
                                          : (const void*)qcdev::root_rows_lds<false, qcdev::ROWS_THREADS>;
    // (the kernel's static LDS, the long-list table, comes on top)
    bool lds_rows = !g->has_n2 && lds_bytes + sizeof(uint32_t) * (qcdev::ROWS_BIG_MAX + 4) <= (size_t)lds_max;
    if (lds_rows && lds_bytes > 48 * 1024 &&
        hipFuncSetAttribute(rows_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
        lds_rows = false;
    if (!lds_rows) {
        QC_HIP(hipMemsetAsync(I.A1.p, 0, mat_bytes, s));
        QC_HIP(hipMemsetAsync(I.A2.p, 0, mat_bytes, s));
    }
    DevCSR* G = nullptr;
    if ((rc = device_csr(g, &G, s))) return rc;
    // temporaries carry the stream: released on an error return while the
    // kernels below still run, they are not handed out again before those end
    DBuf d_loc, d_n2, d_vis;
    if ((rc = d_loc.alloc((size_t)n * 4, s))) return rc;
    if ((rc = d_n2.alloc((size_t)R * 4, s))) return rc;
    if ((rc = I.d_orig.alloc((size_t)R * 4))) return rc;
    QC_HIP(hipMemcpyAsync(d_loc.p, loc.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(I.d_orig.p, I.orig.data(), (size_t)R * 4, hipMemcpyHostToDevice, s));
    int blocks = std::min(R, 2048);
    uint64_t vis_words = ((uint64_t)n + 63) / 64;
    const bool phases = getenv("QCMINE_PHASES") != nullptr;
    auto stamp = [&]() {
        if (phases) (void)hipStreamSynchronize(s);
        return std::chrono::steady_clock::now();
    };
    const auto t_up = stamp();
    if (lds_rows) {
        // one block per root vertex: the dispatcher hands the next vertex to
        // whichever slot frees first (their costs differ by orders of magnitude)
        const int lblocks = std::min(R, 1 << 20);
        const uint64_t* offp = G->off.as<uint64_t>();
        const uint32_t* adjp = G->adj.as<uint32_t>();
        const int32_t* locp = d_loc.as<int32_t>();
        const uint32_t* origp = I.d_orig.as<uint32_t>();
        int32_t Rv = R, rwv = I.rw, nv = n;
        uint64_t* A1p = I.A1.as<uint64_t>();
        uint64_t* A2p = I.A2.as<uint64_t>();
        int32_t* n2p = d_n2.as<int32_t>();
        void* rargs[] = {(void*)&offp, (void*)&adjp, (void*)&locp, (void*)&origp, (void*)&Rv,
                         (void*)&rwv,  (void*)&nv,   (void*)&A1p,  (void*)&A2p,   (void*)&n2p};
        QC_HIP(hipLaunchKernel(rows_kernel, dim3(lblocks), dim3(rows_threads), rargs, lds_bytes, s));
    } else {
        if (!g->has_n2) {
            if ((rc = d_vis.alloc((size_t)blocks * vis_words * 8, s))) return rc;
            QC_HIP(hipMemsetAsync(d_vis.p, 0, (size_t)blocks * vis_words * 8, s));
        }
        hipLaunchKernelGGL(qcdev::root_rows, dim3(blocks), dim3(64), 0, s, G->off.as<uint64_t>(),
                           G->adj.as<uint32_t>(), G->off2.as<uint64_t>(), G->adj2.as<uint32_t>(), d_loc.as<int32_t>(),
                           I.d_orig.as<uint32_t>(), R, I.rw, I.A1.as<uint64_t>(), I.A2.as<uint64_t>(),
                           d_n2.as<int32_t>(), d_vis.as<unsigned long long>(), vis_words, g->has_n2 ? 1 : 0);
    }
    QC_HIP(hipGetLastError());
    const auto t_rows = stamp();
    if ((rc = I.d_d2r.alloc((size_t)R * 4))) return rc;
    if ((rc = I.d_d1r.alloc((size_t)R * 4))) return rc;
    hipLaunchKernelGGL(qcdev::root_d2, dim3(blocks), dim3(64), 0, s, I.A2.as<uint64_t>(), I.A1.as<uint64_t>(), R,
                       I.rw, I.d_d2r.as<int32_t>(), I.d_d1r.as<int32_t>());
    QC_HIP(hipGetLastError());
    I.d2r.resize(R);
    QC_HIP(hipMemcpyAsync(I.d2r.data(), I.d_d2r.p, (size_t)R * 4, hipMemcpyDeviceToHost, s));
    std::vector<int32_t> n2(R);
    QC_HIP(hipMemcpyAsync(n2.data(), d_n2.p, (size_t)R * 4, hipMemcpyDeviceToHost, s));
    if ((rc = I.d_mindeg.alloc(I.mindeg.size() * 4))) return rc;
    if ((rc = I.d_ubq.alloc(I.ubq.size() * 4))) return rc;
    QC_HIP(hipMemcpyAsync(I.d_mindeg.p, I.mindeg.data(), I.mindeg.size() * 4, hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(I.d_ubq.p, I.ubq.data(), I.ubq.size() * 4, hipMemcpyHostToDevice, s));
    QC_HIP(hipStreamSynchronize(s));
    if (phases) {
        const auto t_dev = std::chrono::steady_clock::now();
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        fprintf(stderr,
                "[qcmine] root: R=%d host filter/peel %.2f ms (from call; filter %.2f, degrees %.2f, peel %.2f), "
                "device rows %.2f ms (%s; uploads %.2f, rows %.2f, degrees and tables %.2f)\n",
                R, ms(t_start, t_host), ms(t_start, t_filter), ms(t_filter, t_count), ms(t_count, t_host),
                ms(t_host, t_dev), lds_rows ? "lds" : "global", ms(t_host, t_up), ms(t_up, t_rows), ms(t_rows, t_dev));
    }

    // ---- cover vertex and covered set (main.cu:953-975) ----
    int maxd = 0, maxi = -1;
    for (int a = 0; a < R; a++) {
        int v = (int)I.orig[a];
        if (exd[v] > maxd) {
            maxd = exd[v];
            maxi = a;
        }
    }
    struct RV {
        int32_t id, label, exdeg, l2;
    };
    std::vector<RV> rv(R);
    for (int a = 0; a < R; a++) rv[a] = {a, 0, exd[I.orig[a]], n2[a]};
    if (maxi >= 0) {
        rv[maxi].label = 3;
        int v = (int)I.orig[maxi];
        for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
            int la = loc[g->adj[j]];
            if (la >= 0 && rv[la].label == 0) rv[la].label = 2;
        }
    }
    // ---- Q order (main.cu:2059-2104); indeg is 0 for every root vertex ----
    std::sort(rv.begin(), rv.end(), [](const RV& x, const RV& y) {
        int rx = label_rank(x.label), ry = label_rank(y.label);
        if (rx != ry) return rx < ry;
        if (x.exdeg != y.exdeg) return x.exdeg > y.exdeg;
        if (x.l2 != y.l2) return x.l2 > y.l2;
        return x.id > y.id;
    });
    I.root.resize(R);
    for (int a = 0; a < R; a++) I.root[a] = Vrec{rv[a].id, rv[a].label, 0, rv[a].exdeg};

    I.dg.R = R;
    I.dg.rw = I.rw;
    I.dg.A1 = I.A1.as<uint64_t>();
    I.dg.A2 = I.A2.as<uint64_t>();
    I.dg.mindeg = I.d_mindeg.as<int32_t>();
    I.dg.mindeg_n = (int32_t)I.mindeg.size();
    I.dg.ubq = I.d_ubq.as<int32_t>();
    I.dg.ubq_n = (int32_t)I.ubq.size();
    I.dg.min_size = ms;
    I.dg.gamma = gamma;
    I.dg.orig = I.d_orig.as<uint32_t>();
    I.dg.d2r = I.d_d2r.as<int32_t>();
    I.dg.d1r = I.d_d1r.as<int32_t>();
    I.root_bound = 0;
    I.root_branches = 0;
    for (int a = 0; a < R; a++)
        if (I.root[a].label != 2) I.root_branches++;
    for (int a = 0; a < R; a++) {
        const Vrec& x = I.root[a];
        if (x.label != 1) I.root_bound += std::min<uint64_t>((uint64_t)R, (uint64_t)I.d2r[x.id] + 1);
    }
    return QC_OK;
}

}  // namespace

int build_instance(const qc_graph* g, double gamma, int ms, Instance& I, hipStream_t s) {
    const int rc = build_instance_impl(g, gamma, ms, I, s);
    // an error can leave the root kernels running on the instance's tables:
    // they finish before the caller may release the instance
    if (rc != QC_OK) (void)hipStreamSynchronize(s);
    return rc;
}

int root_window_cut(const Instance& I, long long root_begin, long long root_end, RootCut& c, hipStream_t s) {
    c = RootCut();
    c.b0 = root_begin;
    c.b1 = root_end > 0 ? root_end : -1;
    const int R = I.R;
    if (R == 0) return QC_OK;
    // the window as the expansion resolves it (branches exclude the covered
    // block; a negative start counts from the end)
    const long long nb = (long long)I.root_branches;
    const long long rb0 = root_begin >= 0 ? root_begin : std::max(0ll, nb + root_begin);
    const long long bbeg = std::min(rb0, nb);
    long long bend = nb;
    if (root_end > 0 && root_end < bend) bend = root_end;
    if (bbeg <= 0 || bbeg >= bend) return QC_OK;
    const int ts = R - (int)bbeg;
    // worth a launch only when one wave would scan many row words (kept rows
    // x row words; QCMINE_ROOT_CUT_WORDS overrides the 2^20 threshold)
    long long min_words = 1ll << 20;
    if (const char* e = getenv("QCMINE_ROOT_CUT_WORDS")) min_words = atoll(e);
    if ((long long)ts * I.rw < min_words) return QC_OK;
    // the lookahead of the whole root (main.cu:1349-1405) must fail, since the
    // expansion does not repeat it on the shorter record: every vertex needs
    // |N2 n root| >= R - 1 and indeg + exdeg >= mindeg[R] to pass
    const int req = (int)std::ceil(I.gamma * (double)(R - 1));
    bool passes = true;
    for (const Vrec& v : I.root)
        if (I.d2r[v.id] < R - 1 || v.indeg + v.exdeg < req) {
            passes = false;
            break;
        }
    if (passes) return QC_OK;
    // removed: positions [ts, R) (branches 0 .. bbeg - 1 take them from the end)
    std::vector<uint64_t> mask((size_t)I.rw, 0);
    for (int q = ts; q < R; q++) {
        const uint32_t a = (uint32_t)I.root[q].id;
        mask[a >> 6] |= 1ull << (a & 63);
    }
    std::vector<uint32_t> ids((size_t)ts);
    for (int q = 0; q < ts; q++) ids[q] = (uint32_t)I.root[q].id;
    DBuf d_mask, d_ids, d_cnt;
    int rc;
    if ((rc = d_mask.alloc((size_t)I.rw * 8, s)) || (rc = d_ids.alloc((size_t)ts * 4, s)) ||
        (rc = d_cnt.alloc((size_t)ts * 4, s)))
        return rc;
    QC_HIP(hipMemcpyAsync(d_mask.p, mask.data(), (size_t)I.rw * 8, hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_ids.p, ids.data(), (size_t)ts * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(qcdev::root_cut_deg, dim3(std::min(ts, 8192)), dim3(64), 0, s, I.A1.as<uint64_t>(), I.rw,
                       d_ids.as<uint32_t>(), ts, d_mask.as<uint64_t>(), d_cnt.as<int32_t>());
    QC_HIP(hipGetLastError());
    std::vector<int32_t> cnt((size_t)ts), d1r((size_t)R);
    QC_HIP(hipMemcpyAsync(cnt.data(), d_cnt.p, (size_t)ts * 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipMemcpyAsync(d1r.data(), I.d_d1r.p, (size_t)R * 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    c.rec.assign(I.root.begin(), I.root.begin() + ts);
    for (int q = 0; q < ts; q++) c.rec[q].exdeg -= cnt[q];
    // removal k scans N(r) of root vertex r at position R - k (main.cu:
    // 1437-1450); the lookahead scanned N2 of every root vertex (1373-1381)
    for (int q = ts; q < R; q++) {
        const int a = I.root[q].id;
        c.w_adj += (uint64_t)d1r[a] + (uint64_t)I.d2r[a];
    }
    c.cut = (int32_t)bbeg;
    c.b0 = 0;
    c.b1 = bend - bbeg;
    return QC_OK;
}

}  // namespace qc
