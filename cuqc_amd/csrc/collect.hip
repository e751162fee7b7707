// collect.hip -- the raw sets of a persistent launch, brought to the host.
//
// The expansion kernels append raw sets to one clique stream (a ref
// {off, len, inst} per set, its vertices at off) in whatever order the waves
// finish them.  The reference copies its device clique buffer to the host
// and appends it to temp.txt after the enumeration timer stops
// (dump_cliques, main.cu:688, 1234-1246).  Here the sets are put in order on
// the device -- instances one after another, each instance's sets in the
// order of their vertex slots -- and packed into one contiguous array, so
// the host receives two finished arrays (offsets, vertices) and does no
// sorting or copying of its own:
//   1. key = inst << 40 | off per ref, radix sort of (key, ref index);
//   2. set lengths in that order, exclusive scan -> output offsets;
//   3. one thread per set copies its vertices to its output offset.

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <vector>

#include "qc_internal.h"

namespace qcdev {

__global__ void rc_keys(const CliqueRef* __restrict__ refs, uint64_t k, uint64_t* __restrict__ keys,
                        uint32_t* __restrict__ idx) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    const CliqueRef r = refs[i];
    keys[i] = ((uint64_t)r.inst << 40) | r.off;
    idx[i] = (uint32_t)i;
}

__global__ void rc_lens(const CliqueRef* __restrict__ refs, const uint32_t* __restrict__ sidx, uint64_t k,
                        uint64_t* __restrict__ lens) {
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j < k) lens[j] = refs[sidx[j]].len;
    if (j == k) lens[k] = 0;
}

__global__ __launch_bounds__(256) void rc_gather(const CliqueRef* __restrict__ refs, const uint32_t* __restrict__ sidx,
                                                 const uint64_t* __restrict__ oo, uint64_t k,
                                                 const uint32_t* __restrict__ v, uint32_t* __restrict__ out) {
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= k) return;
    const CliqueRef r = refs[sidx[j]];
    const uint64_t o = oo[j];
    for (uint32_t t = 0; t < r.len; t++) out[o + t] = v[r.off + t];
}

}  // namespace qcdev

namespace qc {

#define QC_RC(x)                                                            \
    do {                                                                    \
        if ((x) != hipSuccess) {                                            \
            set_error("raw set collection: %s failed", #x);                 \
            return QC_EDEVICE;                                              \
        }                                                                   \
    } while (0)

int collect_raw(const CliqueRef* d_refs, const uint32_t* d_verts, uint64_t k, uint64_t nv, std::vector<uint64_t>& off,
                std::vector<uint32_t>& v, hipStream_t s) {
    off.assign(1, 0);
    v.clear();
    if (k == 0) return QC_OK;
    if (k >= 0xFFFFFFFFull || nv >= (1ull << 40)) {
        set_error("raw set collection: %llu sets of %llu vertices exceed the key fields", (unsigned long long)k,
                  (unsigned long long)nv);
        return QC_EOVERFLOW;
    }
    size_t sort_b = 0, scan_b = 0;
    QC_RC(rocprim::radix_sort_pairs(nullptr, sort_b, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                                    (uint32_t*)nullptr, (size_t)k, 0, 56, s));
    QC_RC(rocprim::exclusive_scan(nullptr, scan_b, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t)0, (size_t)k + 1,
                                  rocprim::plus<uint64_t>(), s));
    DBuf keys, skeys, idx, sidx, lens, oo, out, temp;
    int rc;
    if ((rc = keys.alloc(k * 8, s)) || (rc = skeys.alloc(k * 8, s)) || (rc = idx.alloc(k * 4, s)) ||
        (rc = sidx.alloc(k * 4, s)) || (rc = lens.alloc((k + 1) * 8, s)) || (rc = oo.alloc((k + 1) * 8, s)) ||
        (rc = out.alloc(std::max<uint64_t>(nv, 1) * 4, s)) || (rc = temp.alloc(std::max(sort_b, scan_b), s)))
        return rc;
    const unsigned nb = (unsigned)((k + 256) / 256);
    hipLaunchKernelGGL(qcdev::rc_keys, dim3(nb), dim3(256), 0, s, d_refs, k, keys.as<uint64_t>(), idx.as<uint32_t>());
    QC_RC(hipGetLastError());
    QC_RC(rocprim::radix_sort_pairs(temp.p, sort_b, keys.as<uint64_t>(), skeys.as<uint64_t>(), idx.as<uint32_t>(),
                                    sidx.as<uint32_t>(), (size_t)k, 0, 56, s));
    hipLaunchKernelGGL(qcdev::rc_lens, dim3(nb), dim3(256), 0, s, d_refs, sidx.as<uint32_t>(), k, lens.as<uint64_t>());
    QC_RC(hipGetLastError());
    QC_RC(rocprim::exclusive_scan(temp.p, scan_b, lens.as<uint64_t>(), oo.as<uint64_t>(), (uint64_t)0, (size_t)k + 1,
                                  rocprim::plus<uint64_t>(), s));
    hipLaunchKernelGGL(qcdev::rc_gather, dim3(nb), dim3(256), 0, s, d_refs, sidx.as<uint32_t>(), oo.as<uint64_t>(), k,
                       d_verts, out.as<uint32_t>());
    QC_RC(hipGetLastError());
    off.resize(k + 1);
    QC_RC(hipMemcpyAsync(off.data(), oo.p, (k + 1) * 8, hipMemcpyDeviceToHost, s));
    QC_RC(hipStreamSynchronize(s));
    if (off[k] != nv) {
        set_error("raw set collection: the sets hold %llu vertex slots, the stream %llu", (unsigned long long)off[k],
                  (unsigned long long)nv);
        return QC_EDEVICE;
    }
    v.resize(nv);
    QC_RC(hipMemcpyAsync(v.data(), out.p, nv * 4, hipMemcpyDeviceToHost, s));
    QC_RC(hipStreamSynchronize(s));
    return QC_OK;
}

}  // namespace qc
