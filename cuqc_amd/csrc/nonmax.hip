// nonmax.hip -- RemoveNonMax on the device (main.cu:4569-4961).
//
// The reference inserts every raw set (sorted ascending) into a prefix trie,
// marks leaves that are subsets of sets to their right and writes the
// remaining leaves in trie DFS order.  Its output is: each distinct set that
// is not a proper subset of another raw set, ascending, in lexicographic
// order.  This file computes the same result with data-parallel steps:
//   1. segmented radix sort of every set (rocPRIM);
//   2. an order-independent 64-bit hash per set, radix sort by hash, exact
//      comparison inside equal-hash runs -> one representative per distinct
//      set;
//   3. an inverted index vertex -> sets (count, scan, scatter);
//   4. one wave per set: scan the postings of its rarest vertex and test
//      "S subset of T" for every longer T by binary search;
//   5. merge sort of the survivors in lexicographic order (rocPRIM).

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>

#include "qc_internal.h"

namespace qcdev {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
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

__global__ void nm_gather_unique(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ keep,
                                 const uint32_t* __restrict__ pos, uint64_t k, uint32_t* __restrict__ uniq) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < k; p += (uint64_t)gridDim.x * blockDim.x)
        if (keep[p]) uniq[pos[p]] = ids[p];
}

__global__ void nm_count(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                         const uint32_t* __restrict__ uniq, uint32_t nu, uint32_t* __restrict__ cnt) {
    for (uint32_t u = blockIdx.x; u < nu; u += gridDim.x) {
        uint32_t s = uniq[u];
        for (uint64_t j = off[s] + threadIdx.x; j < off[s + 1]; j += blockDim.x) atomicAdd(&cnt[v[j]], 1u);
    }
}

__global__ void nm_scatter(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                           const uint32_t* __restrict__ uniq, uint32_t nu, const uint32_t* __restrict__ start,
                           uint32_t* __restrict__ cursor, uint32_t* __restrict__ post) {
    for (uint32_t u = blockIdx.x; u < nu; u += gridDim.x) {
        uint32_t s = uniq[u];
        for (uint64_t j = off[s] + threadIdx.x; j < off[s + 1]; j += blockDim.x) {
            uint32_t x = v[j];
            uint32_t at = atomicAdd(&cursor[x], 1u);
            post[start[x] + at] = s;
        }
    }
}

// maximal[u] = 1 iff unique set uniq[u] is not a proper subset of another
__global__ __launch_bounds__(64) void nm_subset(const uint64_t* __restrict__ off, const uint32_t* __restrict__ v,
                                                const uint32_t* __restrict__ uniq, uint32_t nu,
                                                const uint32_t* __restrict__ start, const uint32_t* __restrict__ cnt,
                                                const uint32_t* __restrict__ post, uint32_t* __restrict__ maximal) {
    const int lane = threadIdx.x & 63;
    for (uint32_t u = blockIdx.x; u < nu; u += gridDim.x) {
        const uint32_t s = uniq[u];
        const uint64_t sb = off[s], se = off[s + 1];
        const uint64_t ls = se - sb;
        // rarest vertex of s
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
        for (uint32_t q = pb; q < pe && !sub; q++) {
            const uint32_t t = post[q];
            const uint64_t tb = off[t], te = off[t + 1];
            if (t == s || te - tb <= ls) continue;
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
        if (lane == 0) maximal[u] = sub ? 0u : 1u;
    }
}

struct LexLess {
    const uint64_t* off;
    const uint32_t* v;
    __device__ bool operator()(uint32_t a, uint32_t b) const {
        uint64_t ia = off[a], ea = off[a + 1], ib = off[b], eb = off[b + 1];
        while (ia < ea && ib < eb) {
            uint32_t x = v[ia], y = v[ib];
            if (x != y) return x < y;
            ia++;
            ib++;
        }
        return (ea - ia) < (eb - ib) ? true : false;
    }
};

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

int remove_nonmax_device(const std::vector<uint64_t>& off, const std::vector<uint32_t>& v, std::vector<uint64_t>& mo,
                         std::vector<uint32_t>& mv, hipStream_t s) {
    int rc;
    mo.assign(1, 0);
    mv.clear();
    const uint64_t k = off.size() - 1;
    if (k == 0) return QC_OK;
    if (k >= 0xFFFFFFFFull) {
        set_error("too many raw sets (%llu)", (unsigned long long)k);
        return QC_EOVERFLOW;
    }
    const uint64_t nv = v.size();
    uint32_t maxv = 0;
    for (uint32_t x : v) maxv = std::max(maxv, x);
    const uint32_t nvert = maxv + 1;

    DBuf d_off, d_vin, d_v, d_h, d_hs, d_idx, d_ids, d_keep, d_pos, d_uniq, d_cnt, d_start, d_cur, d_post, d_max,
        d_mpos, d_mlist, d_sorted, d_tmp;
    if ((rc = d_off.alloc((k + 1) * 8))) return rc;
    if ((rc = d_vin.alloc(nv * 4))) return rc;
    if ((rc = d_v.alloc(nv * 4))) return rc;
    QC_HIP(hipMemcpyAsync(d_off.p, off.data(), (k + 1) * 8, hipMemcpyHostToDevice, s));
    QC_HIP(hipMemcpyAsync(d_vin.p, v.data(), nv * 4, hipMemcpyHostToDevice, s));

    // 1. sort every set ascending
    size_t tb = 0;
    QC_RP(rocprim::segmented_radix_sort_keys(nullptr, tb, d_vin.as<uint32_t>(), d_v.as<uint32_t>(), (unsigned int)nv,
                                            (unsigned int)k, d_off.as<uint64_t>(), d_off.as<uint64_t>() + 1, 0, 32,
                                            s));
    if ((rc = d_tmp.alloc(tb))) return rc;
    QC_RP(rocprim::segmented_radix_sort_keys(d_tmp.p, tb, d_vin.as<uint32_t>(), d_v.as<uint32_t>(), (unsigned int)nv,
                                            (unsigned int)k, d_off.as<uint64_t>(), d_off.as<uint64_t>() + 1, 0, 32,
                                            s));
    d_vin.release();

    // 2. hash, sort by hash, dedupe
    if ((rc = d_h.alloc(k * 8)) || (rc = d_hs.alloc(k * 8)) || (rc = d_idx.alloc(k * 4)) ||
        (rc = d_ids.alloc(k * 4)) || (rc = d_keep.alloc(k * 4)) || (rc = d_pos.alloc(k * 4)))
        return rc;
    const unsigned gw = (unsigned)std::min<uint64_t>(k, 65536);
    hipLaunchKernelGGL(qcdev::nm_hash, dim3(gw), dim3(64), 0, s, d_off.as<uint64_t>(), d_v.as<uint32_t>(), k,
                       d_h.as<uint64_t>(), d_idx.as<uint32_t>());
    QC_HIP(hipGetLastError());
    tb = 0;
    QC_RP(rocprim::radix_sort_pairs(nullptr, tb, d_h.as<uint64_t>(), d_hs.as<uint64_t>(), d_idx.as<uint32_t>(),
                                    d_ids.as<uint32_t>(), (size_t)k, 0, 64, s));
    if (tb > d_tmp.bytes && (rc = d_tmp.alloc(tb))) return rc;
    QC_RP(rocprim::radix_sort_pairs(d_tmp.p, tb, d_h.as<uint64_t>(), d_hs.as<uint64_t>(), d_idx.as<uint32_t>(),
                                    d_ids.as<uint32_t>(), (size_t)k, 0, 64, s));
    const unsigned gb = (unsigned)std::min<uint64_t>((k + 255) / 256, 16384);
    hipLaunchKernelGGL(qcdev::nm_dedupe, dim3(gb), dim3(256), 0, s, d_off.as<uint64_t>(), d_v.as<uint32_t>(),
                       d_hs.as<uint64_t>(), d_ids.as<uint32_t>(), k, d_keep.as<uint32_t>());
    QC_HIP(hipGetLastError());
    tb = 0;
    QC_RP(rocprim::exclusive_scan(nullptr, tb, d_keep.as<uint32_t>(), d_pos.as<uint32_t>(), 0u, (size_t)k,
                                  rocprim::plus<uint32_t>(), s));
    if (tb > d_tmp.bytes && (rc = d_tmp.alloc(tb))) return rc;
    QC_RP(rocprim::exclusive_scan(d_tmp.p, tb, d_keep.as<uint32_t>(), d_pos.as<uint32_t>(), 0u, (size_t)k,
                                  rocprim::plus<uint32_t>(), s));
    uint32_t last_pos = 0, last_keep = 0;
    QC_HIP(hipMemcpyAsync(&last_pos, d_pos.as<uint32_t>() + (k - 1), 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipMemcpyAsync(&last_keep, d_keep.as<uint32_t>() + (k - 1), 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    const uint32_t nu = last_pos + last_keep;
    if ((rc = d_uniq.alloc((size_t)nu * 4))) return rc;
    hipLaunchKernelGGL(qcdev::nm_gather_unique, dim3(gb), dim3(256), 0, s, d_ids.as<uint32_t>(),
                       d_keep.as<uint32_t>(), d_pos.as<uint32_t>(), k, d_uniq.as<uint32_t>());
    QC_HIP(hipGetLastError());

    // 3. inverted index
    if ((rc = d_cnt.alloc((size_t)nvert * 4)) || (rc = d_start.alloc((size_t)nvert * 4)) ||
        (rc = d_cur.alloc((size_t)nvert * 4)))
        return rc;
    QC_HIP(hipMemsetAsync(d_cnt.p, 0, (size_t)nvert * 4, s));
    QC_HIP(hipMemsetAsync(d_cur.p, 0, (size_t)nvert * 4, s));
    const unsigned gu = (unsigned)std::max<uint32_t>(1, std::min<uint32_t>(nu, 65536));
    hipLaunchKernelGGL(qcdev::nm_count, dim3(gu), dim3(64), 0, s, d_off.as<uint64_t>(), d_v.as<uint32_t>(),
                       d_uniq.as<uint32_t>(), nu, d_cnt.as<uint32_t>());
    QC_HIP(hipGetLastError());
    tb = 0;
    QC_RP(rocprim::exclusive_scan(nullptr, tb, d_cnt.as<uint32_t>(), d_start.as<uint32_t>(), 0u, (size_t)nvert,
                                  rocprim::plus<uint32_t>(), s));
    if (tb > d_tmp.bytes && (rc = d_tmp.alloc(tb))) return rc;
    QC_RP(rocprim::exclusive_scan(d_tmp.p, tb, d_cnt.as<uint32_t>(), d_start.as<uint32_t>(), 0u, (size_t)nvert,
                                  rocprim::plus<uint32_t>(), s));
    uint64_t npost = 0;
    {
        uint32_t a = 0, b = 0;
        QC_HIP(hipMemcpyAsync(&a, d_start.as<uint32_t>() + (nvert - 1), 4, hipMemcpyDeviceToHost, s));
        QC_HIP(hipMemcpyAsync(&b, d_cnt.as<uint32_t>() + (nvert - 1), 4, hipMemcpyDeviceToHost, s));
        QC_HIP(hipStreamSynchronize(s));
        npost = (uint64_t)a + b;
    }
    if ((rc = d_post.alloc(npost * 4))) return rc;
    hipLaunchKernelGGL(qcdev::nm_scatter, dim3(gu), dim3(64), 0, s, d_off.as<uint64_t>(), d_v.as<uint32_t>(),
                       d_uniq.as<uint32_t>(), nu, d_start.as<uint32_t>(), d_cur.as<uint32_t>(),
                       d_post.as<uint32_t>());
    QC_HIP(hipGetLastError());

    // 4. subset test
    if ((rc = d_max.alloc((size_t)nu * 4)) || (rc = d_mpos.alloc((size_t)nu * 4))) return rc;
    hipLaunchKernelGGL(qcdev::nm_subset, dim3(gu), dim3(64), 0, s, d_off.as<uint64_t>(), d_v.as<uint32_t>(),
                       d_uniq.as<uint32_t>(), nu, d_start.as<uint32_t>(), d_cnt.as<uint32_t>(),
                       d_post.as<uint32_t>(), d_max.as<uint32_t>());
    QC_HIP(hipGetLastError());
    tb = 0;
    QC_RP(rocprim::exclusive_scan(nullptr, tb, d_max.as<uint32_t>(), d_mpos.as<uint32_t>(), 0u, (size_t)nu,
                                  rocprim::plus<uint32_t>(), s));
    if (tb > d_tmp.bytes && (rc = d_tmp.alloc(tb))) return rc;
    QC_RP(rocprim::exclusive_scan(d_tmp.p, tb, d_max.as<uint32_t>(), d_mpos.as<uint32_t>(), 0u, (size_t)nu,
                                  rocprim::plus<uint32_t>(), s));
    uint32_t lp = 0, lm = 0;
    QC_HIP(hipMemcpyAsync(&lp, d_mpos.as<uint32_t>() + (nu - 1), 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipMemcpyAsync(&lm, d_max.as<uint32_t>() + (nu - 1), 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    const uint32_t nmx = lp + lm;
    if ((rc = d_mlist.alloc((size_t)nmx * 4)) || (rc = d_sorted.alloc((size_t)nmx * 4))) return rc;
    hipLaunchKernelGGL(qcdev::nm_gather_unique, dim3(gb), dim3(256), 0, s, d_uniq.as<uint32_t>(),
                       d_max.as<uint32_t>(), d_mpos.as<uint32_t>(), (uint64_t)nu, d_mlist.as<uint32_t>());
    QC_HIP(hipGetLastError());

    // 5. lexicographic order
    qcdev::LexLess less{d_off.as<uint64_t>(), d_v.as<uint32_t>()};
    tb = 0;
    QC_RP(rocprim::merge_sort(nullptr, tb, d_mlist.as<uint32_t>(), d_sorted.as<uint32_t>(), (size_t)nmx, less, s));
    if (tb > d_tmp.bytes && (rc = d_tmp.alloc(tb))) return rc;
    QC_RP(rocprim::merge_sort(d_tmp.p, tb, d_mlist.as<uint32_t>(), d_sorted.as<uint32_t>(), (size_t)nmx, less, s));

    std::vector<uint32_t> order(nmx), sv(nv);
    QC_HIP(hipMemcpyAsync(order.data(), d_sorted.p, (size_t)nmx * 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipMemcpyAsync(sv.data(), d_v.p, nv * 4, hipMemcpyDeviceToHost, s));
    QC_HIP(hipStreamSynchronize(s));
    mo.reserve(nmx + 1);
    for (uint32_t x : order) {
        mv.insert(mv.end(), sv.begin() + off[x], sv.begin() + off[x + 1]);
        mo.push_back(mv.size());
    }
    return QC_OK;
}

}  // namespace qc
