// nonmax.hip -- RemoveNonMax on the device (main.cu:4569-4961).
//
// The reference inserts every raw set (sorted ascending) into a prefix trie,
// marks leaves that are subsets of sets to their right and writes the
// remaining leaves in trie DFS order.  Its output is: each distinct set that
// is not a proper subset of another raw set, ascending, in lexicographic
// order.  This file computes the same result with data-parallel steps over
// the raw sets where the expansion kernels left them (no host round trip):
//   0. copy every set into one contiguous array of keys; a vertex v of
//      instance i becomes base[i] + v, so the sets of different instances of a
//      batch share no key and are never compared with each other;
//   1. segmented radix sort of every set (rocPRIM);
//   2. an order-independent 64-bit hash per set, radix sort by hash, exact
//      comparison inside equal-hash runs -> one representative per distinct
//      set;
//   3. an inverted index key -> sets (count, scan, scatter);
//   4. one wave per set: scan the postings of its rarest key and test
//      "S subset of T" for every longer T by binary search;
//   5. merge sort of the survivors in lexicographic key order (rocPRIM): the
//      instances come out one after another, each in lexicographic order;
//   6. the survivors are written out with their instance ids.
// Every buffer is sized by upper bounds known before the first launch (sets,
// vertices, keys), so the sequence has no host synchronization and no
// allocation of its own: it runs in one caller-provided scratch region.

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <mutex>
#include <vector>

#include "qc_internal.h"

namespace qcdev {

constexpr uint32_t NM_NONE = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

template <class T>
__global__ void nm_fill(T* __restrict__ p, uint64_t n, T x) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

// lens[i] = size of raw set i (i < k), lens[k] = 0.  Caller offsets that
// decrease or pass nv raise *err and give an empty set.
__global__ void nm_lens(const CliqueRef* __restrict__ refs, const uint64_t* __restrict__ offs, uint64_t k,
                        uint64_t nv, uint64_t* __restrict__ lens, uint32_t* __restrict__ err) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= k; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t len = 0;
        if (i < k && refs) {
            len = refs[i].len;
        } else if (i < k) {
            const uint64_t a = offs[i], b = offs[i + 1];
            if (b < a || b > nv)
                atomicOr(err, 1u);
            else
                len = b - a;
        }
        lens[i] = len;
    }
}

// offsets past nv (overlapping caller sets) are cut to nv, raising *err:
// every later access stays inside the nv-entry arrays
__global__ void nm_clamp(uint64_t* __restrict__ o, uint64_t k, uint64_t nv, uint32_t* __restrict__ err) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= k; i += (uint64_t)gridDim.x * blockDim.x)
        if (o[i] > nv) {
            o[i] = nv;
            atomicOr(err, 1u);
        }
}

// set i -> keys [o[i], o[i+1]); a vertex outside its instance's range raises
// *err and is clamped into the range (the result is then discarded)
__global__ __launch_bounds__(64) void nm_keys(const CliqueRef* __restrict__ refs, const uint64_t* __restrict__ offs,
                                              const uint32_t* __restrict__ verts, const uint32_t* __restrict__ base,
                                              const uint32_t* __restrict__ limit, uint32_t nkeys, uint64_t k,
                                              const uint64_t* __restrict__ o, uint32_t* __restrict__ keys,
                                              uint32_t* __restrict__ inst, uint32_t* __restrict__ err) {
    const int lane = threadIdx.x & 63;
    for (uint64_t s = blockIdx.x; s < k; s += gridDim.x) {
        const uint64_t src = refs ? refs[s].off : offs[s];
        const uint32_t in = refs ? refs[s].inst : 0u;
        const uint32_t b = base ? base[in] : 0u;
        const uint32_t lim = base ? limit[in] : nkeys;
        const uint64_t dst = o[s], len = o[s + 1] - o[s];
        bool bad = false;
        for (uint64_t j = lane; j < len; j += 64) {
            uint32_t v = verts[src + j];
            if (v >= lim) {
                bad = true;
                v = 0;
            }
            keys[dst + j] = b + v;
        }
        if (__ballot(bad) && lane == 0) atomicOr(err, 1u);
        if (lane == 0) inst[s] = in;
    }
}

__global__ __launch_bounds__(64) void nm_hash(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                                              uint64_t k, uint64_t* __restrict__ h, uint32_t* __restrict__ idx) {
    const int lane = threadIdx.x & 63;
    for (uint64_t s = blockIdx.x; s < k; s += gridDim.x) {
        uint64_t acc = 0;
        for (uint64_t j = off[s] + lane; j < off[s + 1]; j += 64) acc += mix64(v[j]);
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) {
            h[s] = acc ^ mix64(0x51ed270b27a4f5c3ULL + (off[s + 1] - off[s]));
            idx[s] = (uint32_t)s;
        }
    }
}

__device__ __forceinline__ bool same_set(const uint64_t* off, const uint32_t* v, uint32_t a, uint32_t b) {
    uint64_t la = off[a + 1] - off[a], lb = off[b + 1] - off[b];
    if (la != lb) return false;
    for (uint64_t j = 0; j < la; j++)
        if (v[off[a] + j] != v[off[b] + j]) return false;
    return true;
}

// keep[p] = 1 iff the set at sorted position p differs from every earlier set
// of its equal-hash run
__global__ void nm_dedupe(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                          const uint64_t* __restrict__ hs, const uint32_t* __restrict__ ids, uint64_t k,
                          uint32_t* __restrict__ keep) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < k; p += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t me = ids[p];
        uint32_t kp = 1;
        for (uint64_t q = p; q > 0 && hs[q - 1] == hs[p]; q--) {
            if (same_set(off, v, ids[q - 1], me)) {
                kp = 0;
                break;
            }
        }
        keep[p] = kp;
    }
}

// out[pos[p]] = ids[p] for every flagged p < k
__global__ void nm_compact(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ flag,
                           const uint32_t* __restrict__ pos, uint64_t k, uint32_t* __restrict__ out) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < k; p += (uint64_t)gridDim.x * blockDim.x)
        if (flag[p]) out[pos[p]] = ids[p];
}

// *n = number of flagged entries (exclusive scan + last flag)
__global__ void nm_total(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ flag, uint64_t k,
                         uint32_t* __restrict__ n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *n = k ? pos[k - 1] + flag[k - 1] : 0u;
}

__global__ void nm_count(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                         const uint32_t* __restrict__ uniq, const uint32_t* __restrict__ nu_p,
                         uint32_t* __restrict__ cnt) {
    const uint32_t nu = *nu_p;
    for (uint32_t u = blockIdx.x; u < nu; u += gridDim.x) {
        uint32_t s = uniq[u];
        for (uint64_t j = off[s] + threadIdx.x; j < off[s + 1]; j += blockDim.x) atomicAdd(&cnt[v[j]], 1u);
    }
}

__global__ void nm_scatter(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                           const uint32_t* __restrict__ uniq, const uint32_t* __restrict__ nu_p,
                           const uint32_t* __restrict__ start, uint32_t* __restrict__ cursor,
                           uint32_t* __restrict__ post) {
    const uint32_t nu = *nu_p;
    for (uint32_t u = blockIdx.x; u < nu; u += gridDim.x) {
        uint32_t s = uniq[u];
        for (uint64_t j = off[s] + threadIdx.x; j < off[s + 1]; j += blockDim.x) {
            uint32_t x = v[j];
            uint32_t at = atomicAdd(&cursor[x], 1u);
            post[start[x] + at] = s;
        }
    }
}

// sort keys of the postings: the complement of the posted set's size (so
// an ascending sort puts the largest sets first); past the postings ~0
__global__ void nm_post_keys(const uint64_t* __restrict__ off, const uint32_t* __restrict__ post,
                             const uint32_t* __restrict__ start, const uint32_t* __restrict__ cnt, uint32_t nkeys,
                             uint64_t nv, uint32_t* __restrict__ pkey, uint32_t* __restrict__ pend) {
    const uint64_t npost = (uint64_t)start[nkeys - 1] + cnt[nkeys - 1];
    const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nv; q += nt) {
        uint32_t key = 0xFFFFFFFFu;
        if (q < npost) {
            const uint32_t t = post[q];
            key = ~(uint32_t)(off[t + 1] - off[t]);
        }
        pkey[q] = key;
    }
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nkeys; x += nt) pend[x] = start[x] + cnt[x];
}

// maximal[u] = 1 iff unique set uniq[u] (u < nu) is not a proper subset of
// another; 0 for nu <= u < k.  The candidates are the sets posted under the
// rarest key of s that are longer than s; postings list the longest first,
// so the scan ends with the first chunk that reaches a set no longer than s.
// The wave tests 64 candidates at once, one per lane: s sits sorted in LDS,
// and a lane walks its candidate's sorted keys (eight loads in flight) and
// merges them against s; a key of s smaller than the candidate's current key
// is missing from it, which ends the lane's walk.  Sets of more than
// NM_SUB_LDS keys are tested the other way round: the lanes take s's keys and
// binary-search each candidate in turn.
constexpr uint32_t NM_SUB_LDS = 2048;

__global__ __launch_bounds__(64) void nm_subset(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                                                const uint32_t* __restrict__ uniq, const uint32_t* __restrict__ nu_p,
                                                uint64_t k, const uint32_t* __restrict__ start,
                                                const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ post,
                                                uint32_t* __restrict__ maximal) {
    __shared__ uint32_t ss[NM_SUB_LDS];
    const int lane = threadIdx.x & 63;
    const uint32_t nu = *nu_p;
    for (uint64_t u = blockIdx.x; u < k; u += gridDim.x) {
        if (u >= nu) {
            if (lane == 0) maximal[u] = 0u;
            continue;
        }
        const uint32_t s = uniq[u];
        const uint64_t sb = off[s], se = off[s + 1];
        const uint64_t ls = se - sb;
        // rarest key of s
        uint32_t best = 0xFFFFFFFFu, bx = 0;
        for (uint64_t j = sb + lane; j < se; j += 64) {
            uint32_t x = v[j];
            if (cnt[x] < best) {
                best = cnt[x];
                bx = x;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            uint32_t ob = __shfl_xor(best, o, 64), ox = __shfl_xor(bx, o, 64);
            if (ob < best || (ob == best && ox < bx)) {
                best = ob;
                bx = ox;
            }
        }
        bool sub = false;
        const uint32_t pb = start[bx], pe = pb + cnt[bx];
        if (ls <= NM_SUB_LDS) {
            __syncthreads();  // the previous set's readers are done with ss
            for (uint64_t j = lane; j < ls; j += 64) ss[j] = v[sb + j];
            __syncthreads();
            for (uint32_t q0 = pb; q0 < pe && !sub; q0 += 64) {
                const uint32_t q = q0 + (uint32_t)lane;
                uint64_t tb = 0, te = 0;
                if (q < pe) {
                    const uint32_t t = post[q];
                    tb = off[t];
                    te = off[t + 1];
                }
                const bool live = q < pe && te - tb > ls;
                // a set no longer than s in this chunk: none further on is longer
                const bool last = __ballot(q < pe && !live) != 0;
                uint64_t i = 0;
                bool fail = !live;
                for (uint64_t j0 = tb; !fail && i < ls; j0 += 8) {
                    uint32_t e[8];
#pragma unroll
                    for (int x = 0; x < 8; x++) e[x] = j0 + x < te ? v[j0 + x] : 0xFFFFFFFFu;
#pragma unroll
                    for (int x = 0; x < 8; x++) {
                        if (!fail && i < ls) {
                            const uint32_t y = ss[i];
                            if (y < e[x])
                                fail = true;  // y is not in the candidate (past te: e = ~0)
                            else if (y == e[x])
                                i++;
                        }
                    }
                }
                sub = __ballot(!fail && i == ls) != 0;
                if (last) break;
            }
        } else {
            for (uint32_t q = pb; q < pe && !sub; q++) {
                const uint32_t t = post[q];
                const uint64_t tb = off[t], te = off[t + 1];
                if (te - tb <= ls) break;
                bool miss = false;
                for (uint64_t j = sb + lane; j < se; j += 64) {
                    uint32_t x = v[j];
                    uint64_t lo = tb, hi = te;
                    while (lo < hi) {
                        uint64_t mid = (lo + hi) >> 1;
                        if (v[mid] < x)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    if (lo == te || v[lo] != x) miss = true;
                }
                sub = __ballot(miss) == 0;
            }
        }
        if (lane == 0) maximal[u] = sub ? 0u : 1u;
    }
}

// entries at or past *n become NM_NONE (sorted last)
__global__ void nm_pad(uint32_t* __restrict__ list, const uint32_t* __restrict__ n_p, uint64_t k) {
    const uint32_t n = *n_p;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
        if (i >= n) list[i] = NM_NONE;
}

// lexicographic order of the sets named by their index; NM_NONE last
struct LexLess {
    const uint64_t* off;
    const uint32_t* v;
    __device__ bool operator()(uint32_t a, uint32_t b) const {
        if (a == NM_NONE || b == NM_NONE) return a != NM_NONE && b == NM_NONE;
        uint64_t ia = off[a], ea = off[a + 1], ib = off[b], eb = off[b + 1];
        while (ia < ea && ib < eb) {
            uint32_t x = v[ia], y = v[ib];
            if (x != y) return x < y;
            ia++;
            ib++;
        }
        return (ea - ia) < (eb - ib);
    }
};

// olens[r] = size of the r-th survivor (0 past the survivors and at k)
__global__ void nm_out_lens(const uint64_t* __restrict__ off, const uint32_t* __restrict__ sorted,
                            const uint32_t* __restrict__ n_p, uint64_t k, uint64_t* __restrict__ olens) {
    const uint32_t n = *n_p;
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r <= k; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = r < n ? sorted[r] : NM_NONE;
        olens[r] = s == NM_NONE ? 0 : off[s + 1] - off[s];
    }
}

// survivors -> output sets (keys back to vertex ids), their instances and
// the count (~0 when a vertex was out of range)
__global__ __launch_bounds__(64) void nm_out_write(const uint64_t* __restrict__ off, const uint32_t* __restrict__ keys,
                                                   const uint32_t* __restrict__ set_inst,
                                                   const uint32_t* __restrict__ base,
                                                   const uint32_t* __restrict__ sorted,
                                                   const uint32_t* __restrict__ n_p, const uint64_t* __restrict__ oo,
                                                   const uint32_t* __restrict__ err, uint64_t* __restrict__ out_count,
                                                   uint64_t* __restrict__ out_off, uint32_t* __restrict__ out_v,
                                                   uint32_t* __restrict__ out_inst) {
    const int lane = threadIdx.x & 63;
    const uint32_t n = *n_p;
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        const uint32_t s = sorted[r];
        const uint32_t in = set_inst[s];
        const uint32_t b = base ? base[in] : 0u;
        const uint64_t src = off[s], len = off[s + 1] - off[s], dst = oo[r];
        for (uint64_t j = lane; j < len; j += 64) out_v[dst + j] = keys[src + j] - b;
        if (lane == 0) {
            out_off[r] = dst;
            if (out_inst) out_inst[r] = in;
        }
    }
    if (blockIdx.x == 0 && lane == 0) {
        out_off[n] = oo[n];
        *out_count = *err ? ~0ull : (uint64_t)n;
    }
}

}  // namespace qcdev

namespace qc {

#define QC_RP(call)                                                             \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            set_error("rocprim %s failed: %s", #call, hipGetErrorString(e_));   \
            return QC_EDEVICE;                                                  \
        }                                                                       \
    } while (0)

namespace {

// Byte layout of the scratch region for k sets, nv vertices and nkeys keys.
struct NmPlan {
    size_t o, lens, oo, keys_in, keys, inst, h, hs, idx, ids, keep, pos, uniq, cnt, start, cur, post, pkey, pkey2,
        post2, mx, mpos, mlist, sorted, ctr, temp, temp_bytes, total;
};

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

int nm_plan(uint64_t k, uint64_t nv, uint32_t nkeys, NmPlan& P, hipStream_t s) {
    // rocPRIM temporary storage: the largest of every call below
    size_t tb = 0, t;
    const unsigned nvu = (unsigned)nv, ku = (unsigned)k;
    uint32_t* u32 = nullptr;
    uint64_t* u64 = nullptr;
#define NM_Q(call)       \
    do {                 \
        t = 0;           \
        QC_RP(call);     \
        tb = std::max(tb, t); \
    } while (0)
    NM_Q(rocprim::exclusive_scan(nullptr, t, u64, u64, (uint64_t)0, (size_t)k + 1, rocprim::plus<uint64_t>(), s));
    NM_Q(rocprim::segmented_radix_sort_keys(nullptr, t, u32, u32, nvu, ku, u64, u64 + 1, 0, 32, s));
    NM_Q(rocprim::radix_sort_pairs(nullptr, t, u64, u64, u32, u32, (size_t)k, 0, 64, s));
    NM_Q(rocprim::exclusive_scan(nullptr, t, u32, u32, 0u, (size_t)k, rocprim::plus<uint32_t>(), s));
    NM_Q(rocprim::exclusive_scan(nullptr, t, u32, u32, 0u, (size_t)nkeys, rocprim::plus<uint32_t>(), s));
    NM_Q(rocprim::segmented_radix_sort_pairs(nullptr, t, u32, u32, u32, u32, nvu, nkeys, u32, u32 + 1, 0, 32, s));
    qcdev::LexLess less{nullptr, nullptr};
    NM_Q(rocprim::merge_sort(nullptr, t, u32, u32, (size_t)k, less, s));
#undef NM_Q
    size_t b = 0;
    auto take = [&](size_t bytes) {
        const size_t at = b;
        b += up256(std::max<size_t>(bytes, 16));
        return at;
    };
    P.o = take((k + 1) * 8);
    P.lens = take((k + 1) * 8);
    P.oo = take((k + 1) * 8);
    P.keys_in = take(nv * 4);
    P.post = P.keys_in;  // keys_in is dead once the sets are sorted
    P.keys = take(nv * 4);
    P.inst = take(k * 4);
    P.h = take(k * 8);
    P.mlist = P.h;  // the hashes are dead once sorted
    P.sorted = P.h + up256(k * 4);
    P.hs = take(k * 8);
    P.idx = take(k * 4);
    P.ids = take(k * 4);
    P.keep = take(k * 4);
    P.pos = take(k * 4);
    P.uniq = take(k * 4);
    P.cnt = take((size_t)nkeys * 4);
    P.start = take((size_t)nkeys * 4);
    P.cur = take((size_t)nkeys * 4);
    P.pkey = take(nv * 4);
    P.pkey2 = take(nv * 4);
    P.post2 = take(nv * 4);
    P.mx = take(k * 4);
    P.mpos = take(k * 4);
    P.ctr = take(16);
    P.temp_bytes = tb;
    P.temp = take(tb);
    P.total = b;
    return QC_OK;
}

}  // namespace

size_t nonmax_scratch_bytes(uint64_t k, uint64_t nv, uint32_t nkeys, hipStream_t s) {
    NmPlan P;
    if (nm_plan(k, nv, nkeys, P, s)) return 0;
    return P.total;
}

int nonmax_run(const NmInput& in, char* scr, size_t scr_bytes, const NmOutput& out, hipStream_t s) {
    const uint64_t k = in.k, nv = in.nv;
    if (k >= 0xFFFFFFFFull || nv >= 0xFFFFFFFFull) {
        set_error("RemoveNonMax: too many raw sets or vertices (%llu sets, %llu vertices)", (unsigned long long)k,
                  (unsigned long long)nv);
        return QC_EOVERFLOW;
    }
    if (k == 0) {
        hipLaunchKernelGGL(qcdev::nm_fill<uint64_t>, dim3(1), dim3(64), 0, s, out.count, (uint64_t)1, (uint64_t)0);
        hipLaunchKernelGGL(qcdev::nm_fill<uint64_t>, dim3(1), dim3(64), 0, s, out.off, (uint64_t)1, (uint64_t)0);
        QC_HIP(hipGetLastError());
        return QC_OK;
    }
    const uint32_t nkeys = std::max<uint32_t>(in.nkeys, 1);
    NmPlan P;
    int rc;
    if ((rc = nm_plan(k, nv, nkeys, P, s))) return rc;
    if (P.total > scr_bytes) {
        set_error("RemoveNonMax needs %zu bytes of scratch, %zu available", P.total, scr_bytes);
        return QC_EOVERFLOW;
    }
    auto U32 = [&](size_t at) { return (uint32_t*)(scr + at); };
    auto U64 = [&](size_t at) { return (uint64_t*)(scr + at); };
    void* temp = scr + P.temp;
    size_t tb = P.temp_bytes;
    uint32_t* err = U32(P.ctr);
    uint32_t* nu_p = U32(P.ctr) + 1;
    uint32_t* nmx_p = U32(P.ctr) + 2;
    const unsigned gk = (unsigned)std::min<uint64_t>((k + 255) / 256 + 1, 16384);
    const unsigned gw = (unsigned)std::min<uint64_t>(k, 65536);
    const unsigned gn = (unsigned)std::min<uint64_t>(((uint64_t)nkeys + 255) / 256, 16384);

    // 0. contiguous keyed copy
    hipLaunchKernelGGL(qcdev::nm_fill<uint32_t>, dim3(1), dim3(64), 0, s, err, (uint64_t)4, 0u);
    hipLaunchKernelGGL(qcdev::nm_lens, dim3(gk), dim3(256), 0, s, in.refs, in.offsets, k, nv, U64(P.lens), err);
    QC_HIP(hipGetLastError());
    QC_RP(rocprim::exclusive_scan(temp, tb, U64(P.lens), U64(P.o), (uint64_t)0, (size_t)k + 1,
                                  rocprim::plus<uint64_t>(), s));
    hipLaunchKernelGGL(qcdev::nm_clamp, dim3(gk), dim3(256), 0, s, U64(P.o), k, nv, err);
    hipLaunchKernelGGL(qcdev::nm_keys, dim3(gw), dim3(64), 0, s, in.refs, in.offsets, in.verts, in.base, in.limit,
                       nkeys, k, U64(P.o), U32(P.keys_in), U32(P.inst), err);
    QC_HIP(hipGetLastError());

    // 1. sort every set ascending
    tb = P.temp_bytes;
    QC_RP(rocprim::segmented_radix_sort_keys(temp, tb, U32(P.keys_in), U32(P.keys), (unsigned int)nv,
                                            (unsigned int)k, U64(P.o), U64(P.o) + 1, 0, 32, s));
    const uint64_t* off = U64(P.o);
    const uint32_t* keys = U32(P.keys);

    // 2. hash, sort by hash, dedupe
    hipLaunchKernelGGL(qcdev::nm_hash, dim3(gw), dim3(64), 0, s, off, keys, k, U64(P.h), U32(P.idx));
    QC_HIP(hipGetLastError());
    tb = P.temp_bytes;
    QC_RP(rocprim::radix_sort_pairs(temp, tb, U64(P.h), U64(P.hs), U32(P.idx), U32(P.ids), (size_t)k, 0, 64, s));
    hipLaunchKernelGGL(qcdev::nm_dedupe, dim3(gk), dim3(256), 0, s, off, keys, U64(P.hs), U32(P.ids), k,
                       U32(P.keep));
    QC_HIP(hipGetLastError());
    tb = P.temp_bytes;
    QC_RP(rocprim::exclusive_scan(temp, tb, U32(P.keep), U32(P.pos), 0u, (size_t)k, rocprim::plus<uint32_t>(), s));
    hipLaunchKernelGGL(qcdev::nm_total, dim3(1), dim3(64), 0, s, U32(P.pos), U32(P.keep), k, nu_p);
    hipLaunchKernelGGL(qcdev::nm_compact, dim3(gk), dim3(256), 0, s, U32(P.ids), U32(P.keep), U32(P.pos), k,
                       U32(P.uniq));
    QC_HIP(hipGetLastError());

    // 3. inverted index
    hipLaunchKernelGGL(qcdev::nm_fill<uint32_t>, dim3(gn), dim3(256), 0, s, U32(P.cnt), (uint64_t)nkeys, 0u);
    hipLaunchKernelGGL(qcdev::nm_fill<uint32_t>, dim3(gn), dim3(256), 0, s, U32(P.cur), (uint64_t)nkeys, 0u);
    hipLaunchKernelGGL(qcdev::nm_count, dim3(gw), dim3(64), 0, s, off, keys, U32(P.uniq), nu_p, U32(P.cnt));
    QC_HIP(hipGetLastError());
    tb = P.temp_bytes;
    QC_RP(rocprim::exclusive_scan(temp, tb, U32(P.cnt), U32(P.start), 0u, (size_t)nkeys, rocprim::plus<uint32_t>(),
                                  s));
    hipLaunchKernelGGL(qcdev::nm_scatter, dim3(gw), dim3(64), 0, s, off, keys, U32(P.uniq), nu_p, U32(P.start),
                       U32(P.cur), U32(P.post));
    QC_HIP(hipGetLastError());
    // every postings list: largest sets first (the cursor array is free
    // again and takes the segment ends)
    const unsigned gv = (unsigned)std::min<uint64_t>((std::max<uint64_t>(nv, nkeys) + 255) / 256, 16384);
    hipLaunchKernelGGL(qcdev::nm_post_keys, dim3(gv), dim3(256), 0, s, off, U32(P.post), U32(P.start), U32(P.cnt),
                       nkeys, nv, U32(P.pkey), U32(P.cur));
    QC_HIP(hipGetLastError());
    tb = P.temp_bytes;
    QC_RP(rocprim::segmented_radix_sort_pairs(temp, tb, U32(P.pkey), U32(P.pkey2), U32(P.post), U32(P.post2),
                                             (unsigned int)nv, nkeys, U32(P.start), U32(P.cur), 0, 32, s));

    // 4. subset test
    hipLaunchKernelGGL(qcdev::nm_subset, dim3(gw), dim3(64), 0, s, off, keys, U32(P.uniq), nu_p, k, U32(P.start),
                       U32(P.cnt), U32(P.post2), U32(P.mx));
    QC_HIP(hipGetLastError());
    tb = P.temp_bytes;
    QC_RP(rocprim::exclusive_scan(temp, tb, U32(P.mx), U32(P.mpos), 0u, (size_t)k, rocprim::plus<uint32_t>(), s));
    hipLaunchKernelGGL(qcdev::nm_total, dim3(1), dim3(64), 0, s, U32(P.mpos), U32(P.mx), k, nmx_p);
    hipLaunchKernelGGL(qcdev::nm_compact, dim3(gk), dim3(256), 0, s, U32(P.uniq), U32(P.mx), U32(P.mpos), k,
                       U32(P.mlist));
    hipLaunchKernelGGL(qcdev::nm_pad, dim3(gk), dim3(256), 0, s, U32(P.mlist), nmx_p, k);
    QC_HIP(hipGetLastError());

    // 5. lexicographic order (sets past the survivors sort last)
    qcdev::LexLess less{off, keys};
    tb = P.temp_bytes;
    QC_RP(rocprim::merge_sort(temp, tb, U32(P.mlist), U32(P.sorted), (size_t)k, less, s));

    // 6. output
    hipLaunchKernelGGL(qcdev::nm_out_lens, dim3(gk), dim3(256), 0, s, off, U32(P.sorted), nmx_p, k, U64(P.lens));
    QC_HIP(hipGetLastError());
    tb = P.temp_bytes;
    QC_RP(rocprim::exclusive_scan(temp, tb, U64(P.lens), U64(P.oo), (uint64_t)0, (size_t)k + 1,
                                  rocprim::plus<uint64_t>(), s));
    hipLaunchKernelGGL(qcdev::nm_out_write, dim3(gw), dim3(64), 0, s, off, keys, U32(P.inst), in.base,
                       U32(P.sorted), nmx_p, U64(P.oo), err, out.count, out.off, out.v, out.inst);
    QC_HIP(hipGetLastError());
    return QC_OK;
}

int remove_nonmax_sets(const NmInput& in, int n_inst, std::vector<std::vector<uint64_t>*> mo,
                       std::vector<std::vector<uint32_t>*> mv, hipStream_t s) {
    int rc;
    for (int i = 0; i < n_inst; i++) {
        mo[i]->assign(1, 0);
        mv[i]->clear();
    }
    if (in.k == 0) return QC_OK;
    const size_t need = nonmax_scratch_bytes(in.k, in.nv, std::max<uint32_t>(in.nkeys, 1), s);
    if (!need) return QC_EDEVICE;
    DBuf scr, d_count, d_off, d_v, d_inst;
    if ((rc = scr.alloc(need, s)) || (rc = d_count.alloc(8, s)) || (rc = d_off.alloc((in.k + 1) * 8, s)) ||
        (rc = d_v.alloc(in.nv * 4, s)) || (rc = d_inst.alloc(in.k * 4, s)))
        return rc;
    NmOutput out{d_count.as<uint64_t>(), d_off.as<uint64_t>(), d_v.as<uint32_t>(), d_inst.as<uint32_t>()};
    if ((rc = nonmax_run(in, scr.as<char>(), need, out, s))) return rc;
    uint64_t cnt = 0;
    QC_HIP(hipMemcpyAsync(&cnt, out.count, 8, hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    if (cnt == ~0ull) {
        set_error("RemoveNonMax: a raw set holds a vertex outside its graph");
        return QC_EINVAL;
    }
    std::vector<uint64_t> ho(cnt + 1);
    std::vector<uint32_t> hinst(cnt);
    QC_HIP(hipMemcpyAsync(ho.data(), out.off, (cnt + 1) * 8, hipMemcpyDeviceToHost, s));
    QC_HIP(hipMemcpyAsync(hinst.data(), out.inst, cnt * 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> hv(ho[cnt]);
    QC_HIP(hipMemcpyAsync(hv.data(), out.v, ho[cnt] * 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    for (uint64_t r = 0; r < cnt; r++) {
        const uint32_t i = hinst[r];
        if ((int)i >= n_inst) {
            set_error("RemoveNonMax: instance %u out of range", i);
            return QC_EDEVICE;
        }
        mv[i]->insert(mv[i]->end(), hv.begin() + ho[r], hv.begin() + ho[r + 1]);
        mo[i]->push_back(mv[i]->size());
    }
    return QC_OK;
}

int remove_nonmax_device(const std::vector<uint64_t>& off, const std::vector<uint32_t>& v, std::vector<uint64_t>& mo,
                         std::vector<uint32_t>& mv, hipStream_t s) {
    int rc;
    mo.assign(1, 0);
    mv.clear();
    const uint64_t k = off.size() - 1;
    if (k == 0) return QC_OK;
    const uint64_t nv = v.size();
    uint32_t maxv = 0;
    for (uint32_t x : v) maxv = std::max(maxv, x);
    if (maxv == 0xFFFFFFFFu) {
        set_error("vertex id 4294967295 is not supported");
        return QC_EINVAL;
    }
    DBuf d_off, d_v;
    if ((rc = d_off.alloc((k + 1) * 8, s)) || (rc = d_v.alloc(nv * 4, s))) return rc;
    QC_HIP(hipMemcpyAsync(d_off.p, off.data(), (k + 1) * 8, hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_v.p, v.data(), nv * 4, hipMemcpyHostToDevice, s));
    NmInput in{};
    in.k = k;
    in.nv = nv;
    in.offsets = d_off.as<uint64_t>();
    in.verts = d_v.as<uint32_t>();
    in.nkeys = maxv + 1;
    return remove_nonmax_sets(in, 1, {&mo}, {&mv}, s);
}

}  // namespace qc

// ---- RemoveNonMax without allocation ----
// Every buffer of qc_remove_nonmax_noalloc lives in `scratch`, a file-scope
// device array: the call allocates nothing.  Its device address is looked up
// once per device (hipGetSymbolAddress) and kept in scratch_dev[]; the
// array's name itself is not a device address in host code.
namespace {
constexpr size_t NM_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NM_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) char scratch[NM_SCRATCH_BYTES];

namespace qc {
namespace {
std::mutex scratch_mu;
char* scratch_dev[NM_MAX_DEVICES] = {};

int scratch_ptr(char** out) {
    int dev = 0;
    QC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NM_MAX_DEVICES) {
        set_error("device %d: at most %d devices are supported", dev, NM_MAX_DEVICES);
        return QC_EDEVICE;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return QC_EDEVICE;
        }
        scratch_dev[dev] = (char*)p;
    }
    *out = scratch_dev[dev];
    return QC_OK;
}
}  // namespace
}  // namespace qc

extern "C" int qc_remove_nonmax_noalloc(uint64_t count, uint64_t n_verts, const uint64_t* d_offsets,
                                        const uint32_t* d_verts, uint32_t n_vertices, uint64_t* d_out_count,
                                        uint64_t* d_out_offsets, uint32_t* d_out_verts, void* stream) {
    using namespace qc;
    if (!d_out_count || !d_out_offsets || (count && (!d_offsets || !d_verts)) || (n_verts && !d_out_verts)) {
        set_error("NULL argument");
        return QC_EINVAL;
    }
    if (n_vertices == 0 && n_verts) {
        set_error("n_vertices must bound every vertex id");
        return QC_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    char* scr = nullptr;
    int rc = scratch_ptr(&scr);
    if (rc) return rc;
    if (count >= 0xFFFFFFFFull || n_verts >= 0xFFFFFFFFull) {
        set_error("qc_remove_nonmax_noalloc: %llu sets / %llu vertices exceed 2^32 - 1", (unsigned long long)count,
                  (unsigned long long)n_verts);
        return QC_EOVERFLOW;
    }
    const size_t need = count ? nonmax_scratch_bytes(count, n_verts, n_vertices, s) : 0;
    if (count && !need) return QC_EDEVICE;
    if (need > NM_SCRATCH_BYTES) {
        set_error("qc_remove_nonmax_noalloc: %llu sets of %llu vertices (ids < %u) need %zu bytes of scratch; the "
                  "limit is %zu",
                  (unsigned long long)count, (unsigned long long)n_verts, n_vertices, need, NM_SCRATCH_BYTES);
        return QC_EOVERFLOW;
    }
    NmInput in;
    in.k = count;
    in.nv = n_verts;
    in.offsets = d_offsets;
    in.verts = d_verts;
    in.nkeys = n_vertices;
    const NmOutput out{d_out_count, d_out_offsets, d_out_verts, nullptr};
    return nonmax_run(in, scr, NM_SCRATCH_BYTES, out, s);
}
