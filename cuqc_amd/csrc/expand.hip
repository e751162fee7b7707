// expand.hip -- gfx950 expansion of quasi-clique task records.
//
// One 64-lane wavefront owns one task record at a time and runs the whole
// per-task step of the reference host path (h_expand_level body,
// main.cu:1060-1181): lookahead, the sequential branch loop with
// remove-one-vertex, add-one-vertex with diameter and degree pruning,
// critical-vertex pruning, the clique test and the Q-ordered child write.
//
// Adjacency.  Every task is a subset of the root task, so vertex ids are
// root-local (0..R-1) and the 1-hop / 2-hop relations are kept as R x R bit
// matrices in HBM (A1, A2).  A task of t <= TL vertices is staged in LDS
// together with its own t x t sub-matrices (row = position in the task
// record as loaded), so each pruning step is a row AND mask popcount; larger
// tasks run from a per-wave HBM slab and address A1/A2 directly.  Where the
// reference scans adjacency lists through its id->position map
// (vertex_order_map), this code tests bits: the two are the same relation.
//
// Semantics follow the host column of the reference (main.cu:1349-2198),
// including the two degree-update branches of a degree-pruning round
// (recompute from the remaining set when it is smaller than the removed set,
// else decrement via the removed set, main.cu:1775-1807) and the round
// structure that follows critical-vertex pruning when no candidate is
// critical-adjacent: the reference then leaves its id->position map empty
// (main.cu:1575-1602), so a recompute round sets every exdeg to 0 and a
// decrement round changes nothing.  Both effects are reproduced.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qc_device.h"

namespace qcdev {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ int lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Lane broadcasts by v_readlane (lane index uniform): no LDS round trip.
__device__ __forceinline__ int rdlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ unsigned long long rdlane64(unsigned long long x, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ float rdlanef(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// Whole-wave reductions with DPP moves (called with all 64 lanes active).
// Quad permutes and the half-row and row mirrors combine every row of 16
// lanes; row broadcast 15 then adds row 0 into row 1 and row 2 into row 3,
// row broadcast 31 adds rows 0-1 into rows 2-3; lane 63 holds the result.
// Lanes a move does not write keep `id`, the operation's identity.
template <int CTRL, int RM>
__device__ __forceinline__ int dpp32(int id, int v) {
    return __builtin_amdgcn_update_dpp(id, v, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ unsigned long long dpp64(unsigned long long id, unsigned long long v) {
    const uint32_t lo = (uint32_t)dpp32<CTRL, RM>((int)(uint32_t)id, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)dpp32<CTRL, RM>((int)(uint32_t)(id >> 32), (int)(uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
template <class Op>
__device__ __forceinline__ int wave_reduce32(int v, int id, Op op) {
    v = op(v, dpp32<0xb1, 0xf>(id, v));   // quad_perm [1,0,3,2]
    v = op(v, dpp32<0x4e, 0xf>(id, v));   // quad_perm [2,3,0,1]
    v = op(v, dpp32<0x141, 0xf>(id, v));  // row_half_mirror
    v = op(v, dpp32<0x140, 0xf>(id, v));  // row_mirror
    v = op(v, dpp32<0x142, 0xa>(id, v));  // row_bcast:15 into rows 1, 3
    v = op(v, dpp32<0x143, 0xc>(id, v));  // row_bcast:31 into rows 2, 3
    return rdlane(v, WAVE - 1);
}
template <class Op>
__device__ __forceinline__ unsigned long long wave_reduce64(unsigned long long v, unsigned long long id, Op op) {
    v = op(v, dpp64<0xb1, 0xf>(id, v));
    v = op(v, dpp64<0x4e, 0xf>(id, v));
    v = op(v, dpp64<0x141, 0xf>(id, v));
    v = op(v, dpp64<0x140, 0xf>(id, v));
    v = op(v, dpp64<0x142, 0xa>(id, v));
    v = op(v, dpp64<0x143, 0xc>(id, v));
    return rdlane64(v, WAVE - 1);
}
__device__ __forceinline__ int wave_sum(int v) {
    return wave_reduce32(v, 0, [](int x, int y) { return x + y; });
}
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ long long wave_sum64(long long v) {
    return (long long)wave_reduce64((unsigned long long)v, 0ull,
                                    [](unsigned long long x, unsigned long long y) { return x + y; });
}
__device__ __forceinline__ int wave_min(int v) {
    return wave_reduce32(v, INT32_MAX, [](int x, int y) { return min(x, y); });
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
    return wave_reduce64(v, ~0ull, [](unsigned long long x, unsigned long long y) { return x < y ? x : y; });
}
__device__ __forceinline__ int wave_max(int v) {
    return wave_reduce32(v, INT32_MIN, [](int x, int y) { return max(x, y); });
}
// inclusive prefix sum across the wave
__device__ __forceinline__ long long wave_incl_scan64(long long v) {
    int l = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        long long u = __shfl_up(v, o, WAVE);
        if (l >= o) v += u;
    }
    return v;
}

__device__ __forceinline__ void wsync() { __syncthreads(); }

__device__ __forceinline__ bool gbit(const uint64_t* M, int W, int row, int col) {
    return (M[(size_t)row * W + (col >> 6)] >> (col & 63)) & 1ull;
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
struct WS {
    int *pid, *plab, *pin, *pex, *pl2;           // parent record
    int *cid, *cpp, *clab, *cin, *cex, *cl2;     // child under construction
    int *rem, *rmv, *cind, *adjc;                // lists (4 x T2 ints)
    uint64_t *khi, *klo;                         // sort keys, alias the lists
    uint64_t *M1, *M2;                           // local matrices (LDS mode)
    uint64_t *mk0, *mk1;                         // masks (W words each)
    int T2;                                      // capacity rounded to 2^k
};


__device__ inline WS make_ws(char* base, uint32_t tcap, bool mat, uint32_t mwords) {
    WS w;
    uint32_t T2 = pow2_at_least(tcap < 64 ? 64 : tcap);
    uint32_t W = (tcap + 63) / 64;
    int* ip = (int*)base;
    w.pid = ip; ip += T2;
    w.plab = ip; ip += T2;
    w.pin = ip; ip += T2;
    w.pex = ip; ip += T2;
    w.pl2 = ip; ip += T2;
    w.cid = ip; ip += T2;
    w.cpp = ip; ip += T2;
    w.clab = ip; ip += T2;
    w.cin = ip; ip += T2;
    w.cex = ip; ip += T2;
    w.cl2 = ip; ip += T2;
    w.rem = ip; ip += T2;
    w.rmv = ip; ip += T2;
    w.cind = ip; ip += T2;
    w.adjc = ip; ip += T2;
    w.khi = (uint64_t*)w.rem;
    w.klo = (uint64_t*)w.cind;
    size_t b = (size_t)15 * T2 * 4;
    b = (b + 15) & ~(size_t)15;
    uint64_t* up = (uint64_t*)(base + b);
    if (mat) {
        w.M1 = up; up += (size_t)tcap * W;
        w.M2 = up; up += (size_t)tcap * W;
    } else {
        w.M1 = w.M2 = nullptr;
    }
    w.mk0 = up; up += mwords;
    w.mk1 = up; up += mwords;
    w.T2 = (int)T2;
    return w;
}

// per-vertex arrays at `arrays` (LDS), masks of `mwords` words at `masks`
// (root space, in the wave's HBM slab); no local matrices
__device__ inline WS make_ws_split(char* arrays, uint32_t tcap, uint64_t* masks, uint32_t mwords) {
    WS w = make_ws(arrays, tcap, false, 0);
    w.mk0 = masks;
    w.mk1 = masks + mwords;
    return w;
}

// ---------------------------------------------------------------------------
// per-task state
// ---------------------------------------------------------------------------
struct LStats {
    unsigned long long tasks, branches, children, raw, sum_t, sum_tw, sum_clq, large, beyond;
};

// The scalar fields of a DevGraph the kernels read, held in scalar
// registers for the task being expanded: read through the table pointer they
// would be reloaded from memory after every store the compiler cannot
// disambiguate.  The table pointers stay behind `p` (holding all of them in
// scalar registers spills other scalars into vector lanes).
struct GConst {
    const DevGraph* p;
    double gamma;
    int32_t ubq_n, min_size, max_depth;
};

__device__ __forceinline__ int32_t uni32(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ long long uni64(long long x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((unsigned long long)x >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ GConst load_graph(const DevGraph* p) {
    GConst g;
    g.p = p;
    g.gamma = __longlong_as_double(uni64(__double_as_longlong(p->gamma)));
    g.ubq_n = uni32(p->ubq_n);
    g.min_size = uni32(p->min_size);
    g.max_depth = uni32(p->max_depth);
    return g;
}

struct Ctx {
    const ExpandArgs* a;
    GConst g;             // tables of the instance being expanded
    WS w;
    const uint64_t* M1;   // matrix used for tests (local or root space)
    const uint64_t* M2;
    int W;                // words per row of M1/M2 = words per mask
    int lane;
    LStats* st;           // counters of this wave (LDS; lane 0 updates them)
    uint32_t depth;
    uint32_t cap;         // largest task the workspace holds
    uint32_t inst;        // instance of the task being expanded
    long long slot;       // queue slot of the task being expanded (trace)
    unsigned long long* ph;           // phase cycles (LDS, diagnostics)
    long long ph_t;                   // start of the running phase
    long long wadj_l;                 // this lane's share of W_adj (summed at flush)
    bool phase_on;                    // phase counters requested
    bool keys;                        // the task's children are written (inside the depth
                                      // window): only then is the child's lvl2adj kept
    // persistent mode: this wave's private stack (entries [pbot, ptop) of a
    // ring of priv_cap refs; uniform counters, lane 0 touches the entries)
    int shard;                        // this wave's ring shard (blockIdx % q_shards)
    TaskRef* ptopref;                 // LDS: a copy of stack entry ptopi (the last pushed), ~0u: none
    uint32_t ptopi;
    uint32_t* fl;                     // LDS: free block indices of the small classes, FL_LOCAL_MAX per class
    uint32_t* fln;                    // LDS: their counts
    TaskRef* pstack;
    unsigned char* pbkt;
    uint32_t pbot, ptop;
    // children given straight to the rings while waves wait (share_child)
    unsigned long long share_v;       // lane 0: the pending word as read at the last check
    int share_n;                      // children since the last check
    int share_left;                   // children still to give before the next check
    bool share_ok;                    // the running task has SHARE_MIN_T vertices or more
};

// phase accounting: charge the cycles since the last mark to phase k
__device__ __forceinline__ void ph_mark(Ctx& c, int k) {
    if (c.phase_on) {
        long long n = clock64();
        if (c.lane == 0) c.ph[k] += (unsigned long long)(n - c.ph_t);
        c.ph_t = n;
    }
}
__device__ __forceinline__ void ph_start(Ctx& c) {
    if (c.lane == 0)
        for (int k = 0; k < PH_COUNT; k++) c.ph[k] = 0;
    c.ph_t = clock64();
}
__device__ __forceinline__ void ph_flush(Ctx& c) {
    if (c.a->phase && c.lane == 0)
        for (int k = 0; k < PH_COUNT; k++)
            if (c.ph[k]) atomicAdd(&c.a->phase[k], c.ph[k]);
}

// mindeg[i] of the host table (graph_io.cpp build_tables), computed in
// registers: a per-lane table lookup is a dependent memory round trip
__device__ __forceinline__ int md(const Ctx& c, int i) {
    return i < 1 ? 0 : (int)ceil(c.g.gamma * (double)(i - 1));
}
__device__ __forceinline__ int get_mindeg(const Ctx& c, int s) {
    int ms = c.g.min_size;
    return s < ms ? md(c, ms) : md(c, s);
}

// main.cu:2153-2173
__device__ __forceinline__ bool cand_valid(const Ctx& c, int in, int ex, int m, int ub, int lb, int med) {
    int tot = in + ex;
    if (tot < md(c, c.g.min_size)) return false;
    if (tot < get_mindeg(c, m + ex + 1)) return false;
    if (tot < med) return false;
    if (in + ub - 1 < md(c, m + lb)) return false;
    if (tot < get_mindeg(c, m + lb)) return false;
    return true;
}
// main.cu:2175-2198
__device__ __forceinline__ bool vert_ext(const Ctx& c, int in, int ex, int m, int ub, int lb, int med) {
    int tot = in + ex;
    if (tot < md(c, c.g.min_size)) return false;
    if (tot < get_mindeg(c, m + ex)) return false;
    if (tot < med) return false;
    if (ex == 0 && in < get_mindeg(c, m + ex)) return false;
    if (in + ub < md(c, m + ub)) return false;
    if (tot < get_mindeg(c, m + lb)) return false;
    return true;
}

__device__ __forceinline__ int rowpop(const Ctx& c, const uint64_t* M, int row, const uint64_t* mask) {
    const uint64_t* r = M + (size_t)row * c.W;
    int s = 0;
    for (int w = 0; w < c.W; w++) s += __popcll(r[w] & mask[w]);
    return s;
}
// rowpop with four independent loads per step (the gather latency
// overlapped), for the loops that run it over whole records
__device__ __forceinline__ int rowpop4(const Ctx& c, const uint64_t* M, int row, const uint64_t* mask) {
    const uint64_t* r = M + (size_t)row * c.W;
    const int W = c.W;
    int s0 = 0, s1 = 0, w = 0;
    for (; w + 4 <= W; w += 4) {
        const uint64_t x0 = r[w], x1 = r[w + 1], x2 = r[w + 2], x3 = r[w + 3];
        s0 += __popcll(x0 & mask[w]) + __popcll(x2 & mask[w + 2]);
        s1 += __popcll(x1 & mask[w + 1]) + __popcll(x3 & mask[w + 3]);
    }
    for (; w < W; w++) s0 += __popcll(r[w] & mask[w]);
    return s0 + s1;
}

// W_adj bookkeeping: adjacency entries the reference's list scans visit
// (restricted to root vertices), summed over ids.  Each lane keeps its own
// partial sum; the wave sums them when the counters are flushed.
__device__ __forceinline__ void wadj_add(Ctx& c, long long x) {  // x uniform
    if (c.lane == 0) c.wadj_l += x;
}
__device__ __forceinline__ void wadj_lanes(Ctx& c, long long x) {  // x per lane
    c.wadj_l += x;
}

__device__ inline void clear_mask(Ctx& c, uint64_t* mk) {
    for (int w = c.lane; w < c.W; w += WAVE) mk[w] = 0;
    wsync();
}

__device__ inline void set_mask_bit(uint64_t* mk, int b) {
    atomicOr((unsigned long long*)&mk[b >> 6], 1ull << (b & 63));
}

// descending sort of ints a[0..n) (bitonic, padded with INT_MIN)
__device__ void sort_desc_int(Ctx& c, int* a, int n) {
    if (n <= 1) return;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int i = n + c.lane; i < n2; i += WAVE) a[i] = INT32_MIN;
    wsync();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = c.lane; i < n2; i += WAVE) {
                int ixj = i ^ j;
                if (ixj > i) {
                    int x = a[i], y = a[ixj];
                    bool desc = (i & k) == 0;
                    if (desc ? (x < y) : (x > y)) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            wsync();
        }
    }
}

// ascending sort of 128-bit keys (hi, lo) over [0, n)
__device__ void sort_keys(Ctx& c, uint64_t* hi, uint64_t* lo, int n) {
    if (n <= 1) return;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int i = n + c.lane; i < n2; i += WAVE) {
        hi[i] = ~0ull;
        lo[i] = ~0ull;
    }
    wsync();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = c.lane; i < n2; i += WAVE) {
                int ixj = i ^ j;
                if (ixj > i) {
                    uint64_t xh = hi[i], xl = lo[i], yh = hi[ixj], yl = lo[ixj];
                    bool gt = xh > yh || (xh == yh && xl > yl);
                    bool asc = (i & k) == 0;
                    if (asc == gt) {
                        hi[i] = yh;
                        lo[i] = yl;
                        hi[ixj] = xh;
                        lo[ixj] = xl;
                    }
                }
            }
            wsync();
        }
    }
}

// ---------------------------------------------------------------------------
// Candidate indegrees for the LU bounds.  The bounds only need P(k), the sum
// of the k largest indegrees (main.cu:1919-1940 walks them sorted), and
// indegrees lie in [0, m].  With the suffix counts C[v] = #{x >= v},
// P(k) = sum over v >= 1 of min(C[v], k): each of the k largest values x
// adds 1 for every v in [1, x].
//   m < 64: lane v holds C[v] in a register (creg, one ballot per value and
//           chunk); reg = true.
//   else:   tables C[v] (in adjc) and suffix sums S[v] = sum{x : x >= v}
//           (in cind), v in [0, V + 1]; P(k) = S[v + 1] + (k - C[v + 1]) v
//           with v the largest value whose C[v] >= k.
// Returns V = the largest value, or -1 when the tables do not fit (the
// caller sorts).
// ---------------------------------------------------------------------------
// One chunk of candidate indegrees (x < 0: no value in this lane) into the
// register table of the m < 64 case: lane v counts the values >= v, V is the
// largest value.  Loops that produce the indegrees call it on the values in
// their registers, so the table costs no pass of its own.
__device__ __forceinline__ void hist_chunk(const Ctx& c, int x, int& creg, int& V) {
    const int top = uni(wave_max(x));
    V = max(V, top);
    for (int v = 1; v <= top; v++) {
        const int k = __popcll(ballot(x >= v));
        if (c.lane == v) creg += k;
    }
}

__device__ int indeg_dist(Ctx& c, int n, int m, int& creg, bool& reg) {
    WS& w = c.w;
    if (m < WAVE) {
        int cnt = 0, V = 0;
        for (int base = 0; base < n; base += WAVE) {
            const int i = base + c.lane;
            hist_chunk(c, i < n ? w.cind[i] : -1, cnt, V);
        }
        creg = cnt;
        reg = true;
        return V;
    }
    reg = false;
    int mx = 0;
    for (int i = c.lane; i < n; i += WAVE) mx = max(mx, w.cind[i]);
    const int V = uni(wave_max(mx));
    if (V + 2 > w.T2) return -1;
    int* C = w.adjc;
    int* S = w.cind;
    for (int v = c.lane; v <= V + 1; v += WAVE) C[v] = 0;
    wsync();
    for (int i = c.lane; i < n; i += WAVE) atomicAdd(&C[w.cind[i]], 1);
    wsync();
    // suffix scans from the top bin down: C <- counts >= v, S <- sums >= v
    long long cc = 0, cs = 0;
    for (int top = V; top >= 0; top -= WAVE) {
        const int v = top - c.lane;  // lane 0 holds the highest bin of the chunk
        const long long h = v >= 0 ? C[v] : 0;
        const long long ic = wave_incl_scan64(h) + cc;
        const long long is = wave_incl_scan64(h * (v >= 0 ? v : 0)) + cs;
        cc = (long long)rdlane64((unsigned long long)ic, WAVE - 1);
        cs = (long long)rdlane64((unsigned long long)is, WAVE - 1);
        wsync();
        if (v >= 0) {
            C[v] = (int)ic;
            S[v] = (int)is;
        }
        wsync();
    }
    if (c.lane == 0) {
        C[V + 1] = 0;
        S[V + 1] = 0;
    }
    wsync();
    return V;
}

// P(k) = sum_{v=1..V} min(C[v], k) with lane v's register holding C[v]
__device__ __forceinline__ int top_sum_reg(int creg, int V, int k) {
    int p = 0;
    for (int v = 1; v <= V; v++) p += min(__builtin_amdgcn_readlane(creg, v), k);
    return p;
}

// P(k) from the distribution tables (k <= n)
__device__ __forceinline__ long long top_sum(const WS& w, int V, int k) {
    if (k <= 0) return 0;
    const int* C = w.adjc;
    const int* S = w.cind;
    int lo = 0, hi = V;  // C[0] = n >= k
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (C[mid] >= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    return (long long)S[lo + 1] + (long long)(k - C[lo + 1]) * lo;
}

// ---------------------------------------------------------------------------
// LU bounds, main.cu:1857-1972.  Members are child positions [0, m).  With
// V >= 0 the candidate indegrees are given by indeg_dist (creg / reg), with
// V < 0 cind[0..nc) holds them sorted descending.
// ---------------------------------------------------------------------------
__device__ bool lu_bounds(Ctx& c, int m, int nc, int& ub, int& lb, int& med, int V, int creg, bool reg) {
    WS& w = c.w;
    const ExpandArgs& a = *c.a;
    // one pass: the smallest (indeg, exdeg) pair in lexicographic order
    // gives min_in and the smallest exdeg among the members at min_in
    uint64_t min_pair = ((uint64_t)INT32_MAX << 32) | (uint64_t)INT32_MAX;
    int min_tot = INT32_MAX;
    long long sum_in = 0;
    for (int i = c.lane; i < m; i += WAVE) {
        int in = w.cin[i], ex = w.cex[i];
        min_pair = min(min_pair, ((uint64_t)(uint32_t)in << 32) | (uint32_t)ex);
        min_tot = min(min_tot, in + ex);
        sum_in += in;
    }
    min_pair = wave_min_u64(min_pair);
    const int min_in = uni((int)(min_pair >> 32));
    const int min_in_ex = uni((int)(uint32_t)min_pair);
    min_tot = uni(wave_min(min_tot));
    sum_in = wave_sum64(sum_in);

    bool invalid = false;
    med = get_mindeg(c, m + 1);
    if (min_in < md(c, m)) {
        int lb0 = get_mindeg(c, m) - min_in;
        // first x >= lb0 with x > min_in_ex || min_in + x >= mindeg[m + x]
        int x = lb0;
        for (;;) {
            int y = x + c.lane;
            bool stop = (y > min_in_ex) || (min_in + y >= md(c, m + y));
            uint64_t bm = ballot(stop);
            if (bm) {
                x += __builtin_ctzll(bm);
                break;
            }
            x += WAVE;
        }
        lb = x;
        if (min_in + lb < md(c, m + lb)) {
            lb = nc + 1;
            invalid = true;
        }
        int tq = min_tot < c.g.ubq_n ? c.g.p->ubq[min_tot] : c.g.p->ubq[c.g.ubq_n - 1];
        ub = tq + 1 - m;
        if (ub > nc) ub = nc;
        if (lb < ub) {
            // P(idx) = sum of the idx largest candidate indegrees;
            // cond(idx) = sum_in + P(idx) >= m * mindeg[m + idx]
            // i0 = first idx in [lb, ub) with cond, else ub
            long long carry = 0;
            int first = -1, last = -1;
            for (int base = 0; base <= ub; base += WAVE) {
                int idx = base + c.lane;  // P(idx) = prefix over [0, idx)
                long long inc;
                if (V >= 0 && reg) {
                    inc = top_sum_reg(creg, V, idx);
                } else if (V >= 0) {
                    inc = idx <= ub ? top_sum(w, V, idx) : 0;
                } else {
                    long long v = (idx - 1 >= 0 && idx - 1 < nc && idx <= ub) ? (long long)w.cind[idx - 1] : 0;
                    inc = wave_incl_scan64(v) + carry;  // = P(idx)
                    carry = (long long)rdlane64((unsigned long long)inc, WAVE - 1);
                }
                bool inr = idx <= ub;
                bool cnd = inr && (sum_in + inc >= (long long)m * md(c, m + idx));
                uint64_t bfirst = ballot(cnd && idx >= lb);
                if (first < 0 && bfirst) first = base + __builtin_ctzll(bfirst);
                uint64_t blast = ballot(cnd && idx >= lb);
                if (blast) last = base + 63 - __builtin_clzll(blast);
            }
            first = uni(first);
            last = uni(last);
            if (first < 0) {
                lb = ub + 1;
                invalid = true;
            } else {
                // the reference's first loop stops at idx == ub even when
                // cond(ub) is false; first > ub is impossible here
                lb = first;
                int tub = last > first ? last : first;
                if (ub > tub) ub = tub;
                if (lb > 1) med = get_mindeg(c, m + lb);
            }
        }
    } else {
        ub = nc;
        int ms = c.g.min_size;
        lb = m < ms ? ms - m : 0;
    }
    if (m + ub < c.g.min_size) invalid = true;
    if (ub < 0 || ub < lb) invalid = true;
    return invalid;
}

__device__ inline bool members_fail(Ctx& c, int m, int ub, int lb, int med) {
    bool f = false;
    for (int k = c.lane; k < m; k += WAVE)
        if (!vert_ext(c, c.w.cin[k], c.w.cex[k], m, ub, lb, med)) f = true;
    return ballot(f) != 0;
}

// ---------------------------------------------------------------------------
// degree pruning, main.cu:1735-1855.  cind[0..nind) = candidate indegrees.
// ---------------------------------------------------------------------------
// map_empty: the reference's id->position map holds no entry during this call
// (critical-vertex pruning without critical-adjacent candidates).  pre_V >= 0:
// the caller built the register table of cind[0..nind) (pre_V, pre_creg).
__device__ __attribute__((always_inline)) bool degree_pruning(Ctx& c, int& tc, int& nc, int m, int& ub, int& lb,
                                                              int& med, int nind, bool map_empty, int pre_V,
                                                              int pre_creg) {
    WS& w = c.w;
    ph_mark(c, PH_PRUNE);
    int creg = pre_creg;
    bool reg = pre_V >= 0;
    int V = pre_V;
    if (!reg) {
        V = indeg_dist(c, nind, m, creg, reg);
        if (V < 0) sort_desc_int(c, w.cind, nind);
    }
    ph_mark(c, PH_HIST);
    if (lu_bounds(c, m, nind, ub, lb, med, V, creg, reg) || members_fail(c, m, ub, lb, med)) {
        ph_mark(c, PH_LU);
        return true;
    }
    ph_mark(c, PH_LU);
    // classify candidates, stable; first_out = the smallest position ever
    // removed (positions below it stay in place in the compaction)
    int nrem = 0, nrmv = 0, first_out = tc;
    for (int base = m; base < tc; base += WAVE) {
        int i = base + c.lane;
        bool in = i < tc;
        bool ok = in && w.clab[i] == 0 && cand_valid(c, w.cin[i], w.cex[i], m, ub, lb, med);
        uint64_t bo = ballot(ok), bx = ballot(in && !ok);
        if (ok) w.rem[nrem + lanes_below(bo)] = i;
        if (in && !ok) w.rmv[nrmv + lanes_below(bx)] = i;
        if (first_out == tc && bx) first_out = base + __builtin_ctzll(bx);
        nrem += __popcll(bo);
        nrmv += __popcll(bx);
    }
    wsync();
    ph_mark(c, PH_COMPACT);
    while (nrem > 0 && nrmv > 0) {
        // degree update, main.cu:1775-1807
        {
            const bool rc = nrem < nrmv;
            const int nl = rc ? nrem : nrmv;
            const int* L = rc ? w.rem : w.rmv;
            long long sw = 0;
            for (int k = c.lane; k < nl; k += WAVE) sw += c.g.p->d1r[w.cid[L[k]]];
            wadj_lanes(c, sw);
        }
        // only members and remaining candidates are updated: the removed
        // ones never return, so their exdeg is never read again.  The same
        // pass gathers the indegrees of the remaining candidates that are
        // valid under the previous bounds with the new exdeg
        // (main.cu:1811-1815).
        const int nupd = m + nrem;
        const bool rc = nrem < nrmv;
        const int nl = rc ? nrem : nrmv;
        const int* L = rc ? w.rem : w.rmv;
        // 0: exdeg unchanged, 1: all zero (empty map, recompute),
        // 2: bit tests against few vertices, 3: row popcounts against a mask
        const int mode = map_empty ? (rc ? 1 : 0) : (nl < c.W ? 2 : 3);
        if (mode == 2) {
            for (int k = c.lane; k < nl; k += WAVE) w.adjc[k] = w.cpp[L[k]];
            wsync();
        } else if (mode == 3) {
            clear_mask(c, w.mk0);
            for (int k = c.lane; k < nl; k += WAVE) set_mask_bit(w.mk0, w.cpp[L[k]]);
            wsync();
        }
        int nv = 0, hV = 0, hc = 0;
        for (int base = 0; base < nupd; base += WAVE) {
            const int j = base + c.lane;
            bool ok = false;
            int in = 0;
            if (j < nupd) {
                const int i = j < m ? j : w.rem[j - m];
                int ex = w.cex[i];
                if (mode == 1) {
                    ex = 0;
                } else if (mode >= 2) {
                    int cnt = 0;
                    if (mode == 2) {
                        const int row = w.cpp[i];
                        for (int k = 0; k < nl; k++) cnt += (int)gbit(c.M1, c.W, row, w.adjc[k]);
                    } else {
                        cnt = rowpop(c, c.M1, w.cpp[i], w.mk0);
                    }
                    ex = rc ? cnt : ex - cnt;
                }
                if (mode != 0) w.cex[i] = ex;
                if (j >= m) {
                    in = w.cin[i];
                    ok = cand_valid(c, in, ex, m, ub, lb, med);
                }
            }
            const uint64_t bo = ballot(ok);
            // (with m < 64 the register table replaces the list)
            if (ok && m >= WAVE) w.cind[nv + lanes_below(bo)] = in;
            nv += __popcll(bo);
            if (m < WAVE) hist_chunk(c, ok ? in : -1, hc, hV);
        }
        wsync();
        ph_mark(c, PH_DEGUPD);
        ph_mark(c, PH_COMPACT);
        if (m < WAVE) {
            V = hV;
            creg = hc;
            reg = true;
        } else {
            V = indeg_dist(c, nv, m, creg, reg);
            if (V < 0) sort_desc_int(c, w.cind, nv);
        }
        ph_mark(c, PH_HIST);
        if (lu_bounds(c, m, nv, ub, lb, med, V, creg, reg) || members_fail(c, m, ub, lb, med)) {
            ph_mark(c, PH_LU);
            return true;
        }
        ph_mark(c, PH_LU);
        // re-classify the remaining set under the new bounds (in place)
        int kept = 0;
        nrmv = 0;
        for (int base = 0; base < nrem; base += WAVE) {
            int k = base + c.lane;
            bool in = k < nrem;
            int i = in ? w.rem[k] : 0;
            bool ok = in && cand_valid(c, w.cin[i], w.cex[i], m, ub, lb, med);
            uint64_t bo = ballot(ok), bx = ballot(in && !ok);
            if (bx) first_out = min(first_out, rdlane(i, __builtin_ctzll(bx)));
            wsync();
            if (ok) w.rem[kept + lanes_below(bo)] = i;
            if (in && !ok) w.rmv[nrmv + lanes_below(bx)] = i;
            kept += __popcll(bo);
            nrmv += __popcll(bx);
            wsync();
        }
        nrem = kept;
    }
    // compaction: child[m + k] = child[rem[k]] (rem ascending, m + k <= rem[k]);
    // rem[k] = m + k below first_out - m
    ph_mark(c, PH_COMPACT);
    for (int base = first_out - m; base < nrem; base += WAVE) {
        int k = base + c.lane;
        int i = 0, id = 0, pp = 0, lab = 0, in = 0, ex = 0, l2 = 0;
        if (k < nrem) {
            i = w.rem[k];
            id = w.cid[i];
            pp = w.cpp[i];
            lab = w.clab[i];
            in = w.cin[i];
            ex = w.cex[i];
            if (c.keys) l2 = w.cl2[i];
        }
        wsync();
        if (k < nrem) {
            int d = m + k;
            w.cid[d] = id;
            w.cpp[d] = pp;
            w.clab[d] = lab;
            w.cin[d] = in;
            w.cex[d] = ex;
            if (c.keys) w.cl2[d] = l2;
        }
        wsync();
    }
    tc = tc - nc + nrem;
    nc = nrem;
    ph_mark(c, PH_MOVE);
    return false;
}

// ---------------------------------------------------------------------------
// critical vertex pruning, main.cu:1526-1707, up to its closing degree
// pruning: returns 2 when the task fails (the reference's 2), else 0 with the
// arguments of the degree-pruning call that ends the step (candidate
// indegrees in cind[0..nind), map_empty).  The caller runs that call; its
// result gives the reference's 0 / 1.
// ---------------------------------------------------------------------------
__device__ __attribute__((always_inline)) int critical_prep(Ctx& c, int tc, int& nc, int& m, int lb, int& nind_out,
                                                            bool& map_empty) {
    WS& w = c.w;
    int thr = md(c, m + lb);
    // mask of critical members
    // critical members: matrix indices listed in rmv[0..ncm)
    int ncm = 0;
    long long sw = 0;
    for (int base = 0; base < m; base += WAVE) {
        const int k = base + c.lane;
        const bool cr = k < m && w.cin[k] + w.cex[k] == thr && w.cex[k] > 0;
        const uint64_t bc = ballot(cr);
        if (cr) {
            w.rmv[ncm + lanes_below(bc)] = w.cpp[k];
            sw += c.g.p->d1r[w.cid[k]];
        }
        ncm += __popcll(bc);
    }
    wadj_lanes(c, sw);
    wsync();
    // critical-adjacent candidates
    int ncrit = 0;
    if (ncm > 0) {
        const bool sparse = ncm < c.W;
        if (!sparse) {
            clear_mask(c, w.mk0);
            for (int k = c.lane; k < ncm; k += WAVE) set_mask_bit(w.mk0, w.rmv[k]);
            wsync();
        }
        for (int base = m; base < tc; base += WAVE) {
            int i = base + c.lane;
            bool f = false;
            if (i < tc) {
                const int row = w.cpp[i];
                if (sparse) {
                    for (int k = 0; k < ncm && !f; k++) f = gbit(c.M1, c.W, row, w.rmv[k]);
                } else {
                    f = rowpop(c, c.M1, row, w.mk0) > 0;
                }
            }
            ncrit += __popcll(ballot(f));
            if (f) w.clab[i] = 4;
        }
        wsync();
    }
    if (ncrit == 0) {
        // every candidate passes the 2-hop filter; the map stays empty
        int nind = 0;
        for (int base = m; base < tc; base += WAVE) {
            int i = base + c.lane;
            bool in = i < tc;
            uint64_t bo = ballot(in);
            if (in) w.cind[nind + lanes_below(bo)] = w.cin[i];
            nind += __popcll(bo);
        }
        wsync();
        nind_out = nind;
        map_empty = true;
        return 0;
    }
    // stable partition of [m, tc): label 4 first (perm in rmv, temp in cind)
    {
        int nf = 0, nn = 0;
        for (int base = m; base < tc; base += WAVE) {
            int i = base + c.lane;
            bool in = i < tc;
            bool f = in && w.clab[i] == 4;
            uint64_t bf = ballot(f), bn = ballot(in && !f);
            if (f) w.rmv[nf + lanes_below(bf)] = i;
            if (in && !f) w.rmv[ncrit + nn + lanes_below(bn)] = i;
            nf += __popcll(bf);
            nn += __popcll(bn);
        }
        wsync();
        int cnt = tc - m;
        auto permute = [&](int* F) {
            for (int k = c.lane; k < cnt; k += WAVE) w.cind[k] = F[w.rmv[k]];
            wsync();
            for (int k = c.lane; k < cnt; k += WAVE) F[m + k] = w.cind[k];
            wsync();
        };
        permute(w.cid);
        permute(w.cpp);
        permute(w.clab);
        permute(w.cin);
        permute(w.cex);
        if (c.keys) permute(w.cl2);
    }
    // add the critical-adjacent vertices: degree updates and 2-hop counts
    const bool sparse = ncrit < c.W;
    if (!sparse) clear_mask(c, w.mk1);
    sw = 0;
    for (int k = c.lane; k < ncrit; k += WAVE) {
        if (!sparse) set_mask_bit(w.mk1, w.cpp[m + k]);
        sw += c.g.p->d1r[w.cid[m + k]] + c.g.p->d2r[w.cid[m + k]];
    }
    wadj_lanes(c, sw);
    wsync();
    for (int i = c.lane; i < tc; i += WAVE) {
        int pp = w.cpp[i];
        int d1 = 0, d2 = 0;
        if (sparse) {
            for (int k = 0; k < ncrit; k++) {
                const int col = w.cpp[m + k];
                d1 += (int)gbit(c.M1, c.W, pp, col);
                d2 += (int)gbit(c.M2, c.W, pp, col);
            }
        } else {
            d1 = rowpop(c, c.M1, pp, w.mk1);
            d2 = rowpop(c, c.M2, pp, w.mk1);
        }
        w.cin[i] += d1;
        w.cex[i] -= d1;
        w.adjc[i] = d2;
    }
    wsync();
    bool fail = false;
    for (int k = c.lane; k < m; k += WAVE)
        if (w.adjc[k] != ncrit) fail = true;
    for (int k = m + c.lane; k < m + ncrit; k += WAVE)
        if (w.adjc[k] < ncrit - 1) fail = true;
    if (ballot(fail)) return 2;
    for (int k = m + c.lane; k < m + ncrit; k += WAVE) w.clab[k] = 1;
    m += ncrit;
    nc -= ncrit;
    // diameter step around the added vertices
    int nind = 0;
    for (int base = m; base < tc; base += WAVE) {
        int i = base + c.lane;
        bool in = i < tc;
        bool ok = in && w.adjc[i] == ncrit;
        uint64_t bo = ballot(ok);
        if (ok) w.cind[nind + lanes_below(bo)] = w.cin[i];
        if (in && !ok) w.clab[i] = -1;
        nind += __popcll(bo);
    }
    wsync();
    nind_out = nind;
    map_empty = false;
    return 0;
}

// ---------------------------------------------------------------------------
// emission
// ---------------------------------------------------------------------------
__device__ void emit_clique(Ctx& c, const int* ids, int len) {
    const ExpandArgs& a = *c.a;
    unsigned long long ri = 0, si = 0;
    if (c.lane == 0) {
        ri = atomicAdd(&a.clq_ctr->items, 1ull);
        si = atomicAdd(&a.clq_ctr->slots, (unsigned long long)len);
    }
    ri = rdlane64(ri, 0);
    si = rdlane64(si, 0);
    if (ri >= a.clq_ref_cap || si + len > a.clq_vert_cap) {
        if (c.lane == 0) atomicOr(a.error, 1);
        return;
    }
    for (int k = c.lane; k < len; k += WAVE) a.clq_verts[si + k] = c.g.p->orig[ids[k]];
    if (c.lane == 0) {
        CliqueRef r;
        r.off = si;
        r.len = (uint32_t)len;
        r.inst = c.inst;
        a.clq_refs[ri] = r;
    }
    if (c.lane == 0) {
        c.st->raw++;
        c.st->sum_clq += len;
    }
}

// clique test, main.cu:1974-1995
__device__ void check_for_clique(Ctx& c, int m) {
    int req = md(c, m);
    bool bad = false;
    for (int k = c.lane; k < m; k += WAVE)
        if (c.w.cin[k] < req) bad = true;
    if (ballot(bad)) return;
    emit_clique(c, c.w.cid, m);
}

static __device__ __forceinline__ uint32_t label_rank(int l) {
    return l == 1 ? 0u : l == 2 ? 1u : l == 0 ? 2u : l == 3 ? 3u : 4u;
}

// Q order (main.cu:2059-2104) as an ascending 128-bit key
__device__ __forceinline__ void q_key(int lab, int in, int ex, int l2, int id, uint64_t& hi, uint64_t& lo) {
    hi = ((uint64_t)label_rank(lab) << 61) | ((uint64_t)(0x3FFFFFFF - in) << 30) | (uint64_t)(0x3FFFFFFF - ex);
    lo = ((uint64_t)(0x7FFFFFFFu - (uint32_t)l2) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)id);
}

// Cost-model features of a record held in child positions [0, tc) (members
// first): members, candidates, min/max/mean indeg, min/max/mean exdeg, depth,
// ub - lb, mean |N2 n root| over candidates, size.  Means are
// float(sum) / float(count), as in the oracle.
__device__ void task_features(Ctx& c, int tc, int mem, int depth, int gap, float* f /* lane 0 */) {
    WS& w = c.w;
    int mi = INT32_MAX, ma = INT32_MIN, me = INT32_MAX, mx = INT32_MIN;
    long long si = 0, se = 0, s2 = 0;
    for (int i = c.lane; i < tc; i += WAVE) {
        int in = w.cin[i], ex = w.cex[i];
        mi = min(mi, in);
        ma = max(ma, in);
        me = min(me, ex);
        mx = max(mx, ex);
        si += in;
        se += ex;
        if (i >= mem) s2 += c.g.p->d2r[w.cid[i]];
    }
    mi = wave_min(mi);
    ma = wave_max(ma);
    me = wave_min(me);
    mx = wave_max(mx);
    si = wave_sum64(si);
    se = wave_sum64(se);
    s2 = wave_sum64(s2);
    if (c.lane == 0) {
        const int cnd = tc - mem;
        f[0] = (float)mem;
        f[1] = (float)cnd;
        f[2] = (float)mi;
        f[3] = (float)ma;
        f[4] = (float)si / (float)tc;
        f[5] = (float)me;
        f[6] = (float)mx;
        f[7] = (float)se / (float)tc;
        f[8] = (float)depth;
        f[9] = (float)gap;
        f[10] = cnd > 0 ? (float)s2 / (float)cnd : 0.0f;
        f[11] = (float)tc;
    }
}

__device__ __forceinline__ float bf16f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Bucket of a child from the cost model (12 -> 64 -> 64 -> 1 ReLU MLP,
// bf16 weights, fp32 accumulation): lane j evaluates hidden unit j of both
// layers; the prediction is compared with the bucket thresholds.
__device__ int cost_bucket(Ctx& c, const float* feat /* valid in lane 0 */) {
    const ExpandArgs& a = *c.a;
    const uint16_t* W = a.cm_w;
    const float* nrm = a.cm_norm;
    const int j = c.lane;
    float h1 = bf16f(W[12 * 64 + j]);
    for (int k = 0; k < 12; k++) {
        float f = rdlanef(feat[k], 0);
        float z = (log1pf(fmaxf(f, 0.0f)) - nrm[k]) / nrm[12 + k];
        h1 = fmaf(z, bf16f(W[k * 64 + j]), h1);
    }
    h1 = fmaxf(h1, 0.0f);
    const int ow2 = 12 * 64 + 64;
    float h2 = bf16f(W[ow2 + 64 * 64 + j]);
    for (int i = 0; i < 64; i++) h2 = fmaf(rdlanef(h1, i), bf16f(W[ow2 + i * 64 + j]), h2);
    h2 = fmaxf(h2, 0.0f);
    const int ow3 = ow2 + 64 * 64 + 64;
    float y = wave_sum_f(h2 * bf16f(W[ow3 + j])) + bf16f(W[ow3 + 64]);
    int b = 0;
    for (int q = 0; q < a.n_buckets - 1 && q < 3; q++)
        if (y > nrm[26 + q]) b = q + 1;
    return uni(b);
}

// ---------------------------------------------------------------------------
// persistent mode: vertex pool, shared rings, private stacks
// ---------------------------------------------------------------------------
constexpr unsigned long long FL_IDX = (1ull << 36) - 1;  // block index field (units); all ones = empty list

// Free blocks.  Every size class has FL_SHARDS shared lock-free stacks of
// magazines, and a wave uses shard blockIdx % FL_SHARDS (workgroups go to
// the XCDs in turn, so a shard's users share an XCD and its L2).  A
// magazine is a free head block that lists up to FL_LOCAL_MAX - 1 more
// blocks: one exchange on a shard moves up to fl_mag blocks.  The small
// classes (up to 8 units) also keep up to 2 fl_mag block indices per wave
// in LDS (lane 0 only): the free of a task's record and its children's
// allocations meet there without touching memory.
// Head block words: [0] next magazine (index), [1] block count, then the
// other blocks' indices as 32-bit words.
__device__ __forceinline__ unsigned long long* fl_shard(const ExpandArgs& a, int cl, int sh) {
    return a.fl_head + ((size_t)cl * FL_SHARDS + sh) * QSTRIDE;
}

// Pops a magazine (lane 0): its head block index, or FL_IDX when the shard
// holds none.  The next link is read from the head block before the
// exchange; if another wave takes that block meanwhile, the tag in the
// shard head changes and the exchange fails.
__device__ __attribute__((always_inline)) unsigned long long mag_pop(const ExpandArgs& a, unsigned long long* hp) {
    unsigned long long h = __hip_atomic_load(hp, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    while ((h & FL_IDX) != FL_IDX) {
        const unsigned long long idx = h & FL_IDX;
        const unsigned long long nxt =
            __hip_atomic_load((unsigned long long*)(a.pool + idx * POOL_UNIT), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nh = (((h >> 36) + 1ull) << 36) | (nxt & FL_IDX);
        if (__hip_atomic_compare_exchange_strong(hp, &h, nh, __ATOMIC_ACQUIRE, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return idx;
    }
    return FL_IDX;
}

// Pushes the magazine headed by block idx (lane 0; its count and member
// words are written: the release orders them, and the wave's reads of the
// blocks' old records, before any reuse).
__device__ __attribute__((always_inline)) void mag_push(const ExpandArgs& a, unsigned long long* hp, unsigned long long idx) {
    unsigned long long h = __hip_atomic_load(hp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        __hip_atomic_store((unsigned long long*)(a.pool + idx * POOL_UNIT), h & FL_IDX, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nh = (((h >> 36) + 1ull) << 36) | idx;
        if (__hip_atomic_compare_exchange_strong(hp, &h, nh, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

// The shared part of pool_alloc (out of line: it is the rare path, and
// inlined twice it costs the kernel registers).
__device__ __attribute__((noinline)) unsigned long long pool_alloc_shared(const ExpandArgs& a, uint32_t* fl, uint32_t* fln,
                                                                          int cl, uint32_t units) {
    const int sh = (int)(blockIdx.x % FL_SHARDS);
    unsigned long long m = mag_pop(a, fl_shard(a, cl, sh));
    if (m == FL_IDX) {
        const unsigned long long n = (unsigned long long)units * POOL_UNIT;
        const unsigned long long off = atomicAdd(a.pool_top, n);
        if (off + n <= a.pool_cap) return off;
        for (int k = 1; k < FL_SHARDS && m == FL_IDX; k++) m = mag_pop(a, fl_shard(a, cl, (sh + k) % FL_SHARDS));
        if (m == FL_IDX) return ~0ull;
        atomicAdd(a.pool_top, 0ull - n);  // served from a free list after all
    }
    if (cl < FL_SMALL) {
        // the other blocks of the magazine go on the (empty) LDS stack
        const unsigned long long* w = (const unsigned long long*)(a.pool + m * POOL_UNIT);
        const uint32_t cnt = (uint32_t)min<unsigned long long>(max<unsigned long long>(w[1], 1ull), FL_LOCAL_MAX);
        const uint32_t* mem = (const uint32_t*)(w + 2);
        for (uint32_t j = 0; j + 1 < cnt; j++) fl[cl * FL_LOCAL_MAX + j] = mem[j];
        fln[cl] = cnt - 1;
    }
    return m * POOL_UNIT;
}

// A block of the record's size class (lane 0): this wave's LDS stack, a
// magazine of the wave's shard, fresh units from the bump pointer, and once
// the bump region is used up, a magazine of any other shard.  Returns the
// Vrec offset, or ~0 when the pool is exhausted (the bump pointer then keeps
// the failed request: the host reads what the launch needed).
__device__ __attribute__((always_inline)) unsigned long long pool_alloc(Ctx& c, uint32_t nvrec) {
    const ExpandArgs& a = *c.a;
    uint32_t units = 0;
    const int cl = pool_class(nvrec, units);
    const bool small = cl < FL_SMALL;
    if (small) {
        const uint32_t n = c.fln[cl];
        if (n > 0) {
            c.fln[cl] = n - 1;
            return (unsigned long long)c.fl[cl * FL_LOCAL_MAX + n - 1] * POOL_UNIT;
        }
    }
    return pool_alloc_shared(a, c.fl, c.fln, cl, units);
}

// The shared part of pool_free (out of line, like pool_alloc_shared).
__device__ __attribute__((noinline)) void pool_free_shared(const ExpandArgs& a, uint32_t* fl, uint32_t* fln, int cl,
                                                           unsigned long long idx) {
    unsigned long long* w = (unsigned long long*)(a.pool + idx * POOL_UNIT);
    if (cl < FL_SMALL) {
        const uint32_t n = fln[cl];
        const uint32_t k = a.fl_mag - 1;
        uint32_t* mem = (uint32_t*)(w + 2);
        for (uint32_t j = 0; j < k; j++) mem[j] = fl[cl * FL_LOCAL_MAX + n - k + j];
        fln[cl] = n - k;
        w[1] = k + 1;
    } else {
        w[1] = 1;
    }
    mag_push(a, fl_shard(a, cl, (int)(blockIdx.x % FL_SHARDS)), idx);
}

// Gives a block back (lane 0, after the wave has read the record).  A full
// LDS stack hands the block, with the stack's fl_mag - 1 newest entries
// listed in it, to the shard as one magazine.
__device__ __attribute__((always_inline)) void pool_free(Ctx& c, unsigned long long off, uint32_t nvrec) {
    const ExpandArgs& a = *c.a;
    uint32_t units = 0;
    const int cl = pool_class(nvrec, units);
    const unsigned long long idx = off / POOL_UNIT;
    if (cl < FL_SMALL) {
        const uint32_t n = c.fln[cl];
        if (n < 2 * a.fl_mag) {
            c.fl[cl * FL_LOCAL_MAX + n] = (uint32_t)idx;
            c.fln[cl] = n + 1;
            return;
        }
    }
    pool_free_shared(a, c.fl, c.fln, cl, idx);
}

// Claims up to n (<= 64) consecutive free cells of ring b for producers
// (whole wave; lane j checks the cell of position pos + j).  Returns how
// many were claimed (0: the ring is full, or max_tries attempts lost to
// other producers), their first position in *pos.  With a trace the ring
// must not wrap (positions index the trace rows).
__device__ __attribute__((always_inline)) int ring_claim(Ctx& c, int b, int n, unsigned long long* pos_out,
                                                        int max_tries = 1 << 16) {
    const ExpandArgs& a = *c.a;
    unsigned long long* tp = &a.q_tail[b * QSTRIDE];
    const unsigned long long* seq = a.q_seq + (size_t)b * a.q_cap;
    for (int tries = 0; tries < max_tries; tries++) {
        unsigned long long pos = 0;
        if (c.lane == 0) pos = __hip_atomic_load(tp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pos = rdlane64(pos, 0);
        const unsigned long long p = pos + (unsigned long long)c.lane;
        unsigned long long sq = 0;
        if (c.lane < n) sq = __hip_atomic_load(&seq[p % a.q_cap], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t freem = ballot(c.lane < n && sq == p);
        const uint64_t fullm = ballot(c.lane < n && sq < p);
        int k = freem == ~0ull ? WAVE : __builtin_ctzll(~freem);
        k = min(k, n);
        if (a.trace_feat && pos + (unsigned long long)k > a.q_cap) k = (int)(pos < a.q_cap ? a.q_cap - pos : 0);
        if (k == 0) {
            // lane 0's cell still holds the entry of the previous lap: full;
            // otherwise the tail moved under us
            if ((fullm & 1ull) || (a.trace_feat && pos >= a.q_cap)) return 0;
            continue;
        }
        int ok = 0;
        if (c.lane == 0) {
            unsigned long long e = pos;
            ok = __hip_atomic_compare_exchange_strong(tp, &e, pos + (unsigned long long)k, __ATOMIC_RELAXED,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     ? 1
                     : 0;
        }
        if (rdlane(ok, 0)) {
            *pos_out = pos;
            return k;
        }
    }
    return 0;
}

// Publishes positions [pos, pos + k) of bucket b (their refs are written):
// one release fence for the wave's record and ref stores, then the
// sequence words (lanes < k).
__device__ __forceinline__ void ring_publish(Ctx& c, int b, unsigned long long pos, int k) {
    const ExpandArgs& a = *c.a;
    wsync();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (c.lane < k) {
        const unsigned long long p = pos + (unsigned long long)c.lane;
        __hip_atomic_store(&a.q_seq[(size_t)b * a.q_cap + p % a.q_cap], p + 1ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Dynamic order: whether this child goes to a ring instead of the wave's own
// stack.  The stack's records are handed over only between tasks, so while a
// long task runs (a root, a large first-level task) its children would wait
// on this wave while other waves have nothing.  In tasks of SHARE_MIN_T
// vertices or more, every SHARE_EVERY children the wave compares the waves
// that wait for work with the tasks the rings hold (the pending word, as the
// hand-over between tasks does) and gives up to 16 of its next children to
// the rings when they fall short.  Smaller tasks end soon enough for the
// hand-over between tasks, and sending their children through the rings
// costs more than the wait (trees of 30-vertex tasks ran slower).  The word
// is loaded at one check and read at the next, so its latency overlaps the
// children built in between; a count one interval old only delays the answer.
constexpr int SHARE_EVERY = 8;
constexpr uint32_t SHARE_MIN_T = 64;
__device__ __forceinline__ bool share_child(Ctx& c) {
    const ExpandArgs& a = *c.a;
    if (!c.share_ok) return false;
    if (c.share_left > 0) {
        c.share_left--;
        return true;
    }
    if (++c.share_n < SHARE_EVERY) return false;
    c.share_n = 0;
    const unsigned long long ph = rdlane64(c.share_v, 0);
    if (c.lane == 0) c.share_v = __hip_atomic_load(a.pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long hg = (long long)(ph >> 32), pd = (long long)(ph & 0xFFFFFFFFull);
    const long long waiting = pd - ((long long)gridDim.x - hg);
    const long long need = hg - (waiting > 0 ? waiting : 0);
    if (need <= 0) return false;
    c.share_left = (int)min<long long>(need, 16ll) - 1;
    return true;
}

// One task ref into a ring (shared = true first; counted in `pending` before
// anyone can take it) or this wave's private stack (covered by the wave's
// own count while it is busy).  Returns false when both are full (the launch
// then reports a capacity overflow).
__device__ __attribute__((always_inline)) bool push_ref(Ctx& c, const TaskRef& r, int bkt, bool shared, const float* feat) {
    const ExpandArgs& a = *c.a;
    const int rg = bkt * a.q_shards + c.shard;
    for (int pass = 0; pass < 2; pass++) {
        const bool ring = (pass == 0) == (shared || !a.use_priv);
        if (ring) {
            unsigned long long pos = 0;
            if (ring_claim(c, rg, 1, &pos) == 1) {
                const unsigned long long gi = (unsigned long long)rg * a.q_cap + pos % a.q_cap;
                if (c.lane == 0) {
                    __hip_atomic_fetch_add(a.pending, 1ull, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    a.q_refs[gi] = r;
                    if (a.trace_parent) a.trace_parent[gi] = r.b1 > 0 ? -2 : c.slot;
                    if (a.trace_feat && feat)
                        for (int k = 0; k < 12; k++) a.trace_feat[gi * 12 + k] = feat[k];
                }
                ring_publish(c, rg, pos, 1);
                return true;
            }
        } else if (a.use_priv && c.ptop - c.pbot < a.priv_cap) {
            if (c.lane == 0) {
                c.pstack[c.ptop % a.priv_cap] = r;
                c.pbkt[c.ptop % a.priv_cap] = (unsigned char)bkt;
                *c.ptopref = r;  // the entry the next pop takes, kept in LDS too
            }
            c.ptopi = c.ptop;
            c.ptop++;
            return true;
        }
    }
    if (c.lane == 0) {
        atomicAdd(a.overflow, 1ull);
        atomicOr(a.error, 1);
    }
    return false;
}

// Static order: moves the staged children on this wave's stack into the
// ring, up to 64 per claim, oldest first (lane j copies entry j; the
// barrier makes lane 0's earlier stack stores visible to the other lanes).
// Returns false when the ring took none (it is full): the wave then runs
// its staged tasks itself.
__device__ __attribute__((always_inline)) bool static_flush(Ctx& c) {
    const ExpandArgs& a = *c.a;
    bool moved = false;
    wsync();
    while (c.ptop != c.pbot) {
        unsigned long long pos = 0;
        const int k = ring_claim(c, 0, (int)min<uint32_t>(c.ptop - c.pbot, (uint32_t)WAVE), &pos);
        if (k == 0) break;
        if (c.lane < k)
            a.q_refs[(pos + c.lane) % a.q_cap] = c.pstack[(c.pbot + (uint32_t)c.lane) % a.priv_cap];
        if (c.lane == 0)
            __hip_atomic_fetch_add(a.pending, (unsigned long long)k, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        ring_publish(c, 0, pos, k);
        c.pbot += (uint32_t)k;
        moved = true;
    }
    return moved;
}

// TaskRef as eight dwords, broadcast from lane l
__device__ __forceinline__ TaskRef bcast_ref(const TaskRef& mine, int l) {
    uint32_t w[8];
    __builtin_memcpy(w, &mine, sizeof(w));
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = (uint32_t)__builtin_amdgcn_readlane((int)w[k], l);
    TaskRef r;
    __builtin_memcpy(&r, w, sizeof(r));
    return r;
}

__device__ __attribute__((always_inline)) void write_child(Ctx& c, int tc, int mem, int gap) {
    const ExpandArgs& a = *c.a;
    WS& w = c.w;
    if (c.g.max_depth >= 0 && (int)c.depth + 1 > c.g.max_depth) {
        // outside the depth window: counted, never expanded
        if (c.lane == 0) {
            c.st->children++;
            c.st->sum_tw += tc;
            c.st->beyond++;
            if (a.trace_beyond && c.slot >= 0) atomicAdd(&a.trace_beyond[c.slot], 1u);
        }
        return;
    }
    for (int i = c.lane; i < tc; i += WAVE) q_key(w.clab[i], w.cin[i], w.cex[i], w.cl2[i], w.cid[i], w.khi[i], w.klo[i]);
    wsync();
    sort_keys(c, w.khi, w.klo, tc);
    const bool q = a.queue_mode != 0;
    int s = (uint32_t)tc <= a.small_max ? 0 : 1;
    int bucket = 0;
    float feat[12];
    // cost order: children that fit the LDS tile take the lowest bucket
    // without the model (their subtrees are small, and on trees of tiny
    // tasks the features and the forward pass would cost as much as the task)
    const bool predict = q && a.n_buckets > 1 && (uint32_t)tc > a.cost_min_t;
    if (q && (a.trace_feat || predict)) task_features(c, tc, mem, (int)c.depth + 1, gap, feat);
    if (predict) bucket = cost_bucket(c, feat);
    unsigned long long ri = 0, si = 0;
    if (q) ph_mark(c, PH_CHILD);
    if (c.lane == 0) {
        if (q) {
            si = pool_alloc(c, (uint32_t)tc);
        } else {
            ri = atomicAdd(&a.out_ctr[s].items, 1ull);
            si = atomicAdd(&a.out_ctr[s].slots, (unsigned long long)tc);
        }
    }
    ri = rdlane64(ri, 0);
    si = rdlane64(si, 0);
    if (q ? si == ~0ull : (ri >= a.out_ref_cap[s] || si + tc > a.out_vert_cap[s])) {
        if (c.lane == 0) atomicOr(a.error, 1);
        return;
    }
    Vrec* dst = (q ? a.pool : a.out_verts[s]) + si;
    for (int i = c.lane; i < tc; i += WAVE) {
        uint64_t hi = w.khi[i], lo = w.klo[i];
        uint32_t rank = (uint32_t)(hi >> 61);
        Vrec v;
        v.id = (int32_t)(0xFFFFFFFFu - (uint32_t)lo);
        v.label = rank == 0 ? 1 : rank == 1 ? 2 : rank == 2 ? 0 : rank == 3 ? 3 : -1;
        v.indeg = 0x3FFFFFFF - (int32_t)((hi >> 30) & 0x3FFFFFFFull);
        v.exdeg = 0x3FFFFFFF - (int32_t)(hi & 0x3FFFFFFFull);
        dst[i] = v;
    }
    TaskRef r;
    r.off = si;
    r.t = (uint32_t)tc;
    r.depth = (uint16_t)(c.depth + 1);
    r.inst = (uint16_t)c.inst;
    r.bound = 0;
    r.b0 = 0;
    r.b1 = 0;
    if (q) {
        // dynamic order: children stay on this wave's stack (depth first),
        // with their predicted bucket when a cost model orders the work;
        // static order: the stack only stages them, and the wave moves them
        // to the ring (whose positions are dealt to the waves in turn) in
        // one claim when the task ends
        const bool shared = !a.use_priv || (!a.static_sched && share_child(c));
        const bool pushed = push_ref(c, r, bucket, shared, feat);
        ph_mark(c, PH_POOL);
        if (!pushed) return;
    } else {
        // slot bound of this child's own children (see TaskRef)
        unsigned long long bnd = 0;
        for (int i = mem + c.lane; i < tc; i += WAVE) {
            unsigned long long x =
                (unsigned long long)c.g.p->d2r[(int32_t)(0xFFFFFFFFu - (uint32_t)w.klo[i])] + mem + 1;
            bnd += x < (unsigned long long)tc ? x : (unsigned long long)tc;
        }
        bnd = (unsigned long long)wave_sum64((long long)bnd);
        if (c.lane == 0) {
            r.bound = bnd;
            a.out_refs[s][ri] = r;
        }
    }
    if (c.lane == 0) {
        c.st->children++;
        c.st->sum_tw += tc;
    }
}

// first index in [from, t) where pred holds, else t
template <class P>
__device__ __forceinline__ int first_where(Ctx& c, int from, int t, P pred) {
    for (int base = from; base < t; base += WAVE) {
        int i = base + c.lane;
        uint64_t b = ballot(i < t && pred(i));
        if (b) return base + __builtin_ctzll(b);
    }
    return t;
}

// ---------------------------------------------------------------------------
// one task record, main.cu:1060-1181
// ---------------------------------------------------------------------------
template <bool LOCAL>
__device__ __attribute__((always_inline)) void process_task(Ctx& c, const TaskRef ref) {
    const ExpandArgs& a = *c.a;
    WS& w = c.w;
    const int t0 = (int)ref.t;
    c.depth = ref.depth;
    c.keys = !(c.g.max_depth >= 0 && (int)c.depth + 1 > c.g.max_depth);
    const bool item = ref.b1 > 0;
    if (!item) {  // a split item's task is counted by its prepare step
        if (c.lane == 0) {
            c.st->tasks++;
            c.st->sum_t += t0;
            if ((uint32_t)t0 > c.a->small_max) c.st->large++;
        }
    }

    for (int i = c.lane; i < t0; i += WAVE) {
        Vrec v = a.in_verts[ref.off + i];
        w.pid[i] = v.id;
        w.plab[i] = item ? (v.label & 15) : v.label;
        w.pin[i] = v.indeg;
        w.pex[i] = v.exdeg;
        w.pl2[i] = item ? (v.label >> 4) : 0;
    }
    wsync();
    const int nm = uni(first_where(c, 0, t0, [&](int i) { return w.plab[i] != 1; }));
    const int cov = uni(first_where(c, nm, t0, [&](int i) { return w.plab[i] != 2; })) - nm;

    // position -> matrix index: local position (LOCAL) or root id
    int* P2M = LOCAL ? nullptr : w.pid;
    if (LOCAL) {
        // t0 x t0 sub-matrices of A1 / A2, rows in position order
        const int W = (t0 + 63) / 64;
        c.W = W;
        c.M1 = w.M1;
        c.M2 = w.M2;
        const int rw = c.g.p->rw;
        const uint64_t* __restrict__ A1 = c.g.p->A1;
        const uint64_t* __restrict__ A2 = c.g.p->A2;
        // lane = column; the rows' words holding the columns' bits are
        // gathered RU rows at a time (2 RU loads in flight per lane: tasks
        // here are small, and the build is a few rounds of memory latency);
        // the last group masks the rows past t0
        constexpr int RU = 8;
        for (int wd = 0; wd < W; wd++) {
            int col = wd * 64 + c.lane;
            bool cv = col < t0;
            int cb = cv ? w.pid[col] : 0;
            int cw = cb >> 6;
            uint64_t cm = 1ull << (cb & 63);
            for (int i = 0; i < t0; i += RU) {
                uint64_t x[RU], y[RU];
#pragma unroll
                for (int q = 0; q < RU; q++) {
                    x[q] = 0;
                    y[q] = 0;
                    if (i + q < t0) {
                        const size_t rb = (size_t)w.pid[i + q] * rw + cw;
                        x[q] = A1[rb];
                        y[q] = A2[rb];
                    }
                }
#pragma unroll
                for (int q = 0; q < RU; q++) {
                    const uint64_t b = ballot(cv && (x[q] & cm)), e = ballot(cv && (y[q] & cm));
                    if (c.lane == 0 && i + q < t0) {
                        w.M1[(size_t)(i + q) * W + wd] = b;
                        w.M2[(size_t)(i + q) * W + wd] = e;
                    }
                }
            }
        }
        wsync();
    } else {
        c.W = c.g.p->rw;
        c.M1 = c.g.p->A1;
        c.M2 = c.g.p->A2;
    }
#define PM(i) (LOCAL ? (i) : P2M[(i)])
    ph_mark(c, PH_LOAD);

    // ---- lookahead, main.cu:1349-1405 ----
    int nc = t0 - nm;
    if (!item) {
        int req = md(c, t0);
        bool f = false;
        for (int i = c.lane; i < nm; i += WAVE)
            if (w.pin[i] + w.pex[i] < req) f = true;
        if (ballot(f) == 0) {
            // a record holding every root vertex and no member is the whole
            // root: |N2(v) n candidates| is then |N2(v) n root| = d2r[v],
            // with no row scans (O(R^2 / 64) words at the root otherwise).
            // A root shortened by a window's leading removals (root_cut) takes
            // the whole root's values too: its lookahead ran on the whole
            // root, where it failed (the host checked), so nothing is emitted
            const bool cut_root = nm == 0 && ref.depth == 0 && c.g.p->root_cut > 0;
            const bool whole_root = nm == 0 && (t0 == c.g.p->R || cut_root);
            if (!whole_root) clear_mask(c, w.mk0);
            long long sw = 0;
            for (int i = nm + c.lane; i < t0; i += WAVE) {
                if (!whole_root) set_mask_bit(w.mk0, PM(i));
                sw += c.g.p->d2r[w.pid[i]];
            }
            wadj_lanes(c, sw);
            wsync();
            bool bad = false;
            for (int i = nm + c.lane; i < t0; i += WAVE) {
                int l2 = whole_root ? c.g.p->d2r[w.pid[i]] : rowpop(c, c.M2, PM(i), w.mk0);
                w.pl2[i] = l2;
                if (l2 < nc - 1 || w.pin[i] + w.pex[i] < req) bad = true;
            }
            wsync();
            if (ballot(bad) == 0 && !cut_root) {
                emit_clique(c, w.pid, t0);
                ph_mark(c, PH_LOOKAHEAD);
                return;
            }
        }
    }
    ph_mark(c, PH_LOOKAHEAD);

    // ---- branch loop, main.cu:1094-1180 ----
    // Branch ordinal b = j - covered.  Branch b >= 1 first removes the last
    // parent vertex (remove_one_vertex); a failing member ends the loop.
    int tp = t0;
    const int expansions = nc;
    const bool root = ref.depth == 0;
    const int nb = expansions - cov;
    int bend = nb;
    if (root && c.g.p->root_b1 >= 0 && c.g.p->root_b1 < bend) bend = (int)c.g.p->root_b1;
    // a negative window start counts from the end of the root's branches
    const long long rb0 = c.g.p->root_b0 >= 0 ? c.g.p->root_b0 : max(0ll, (long long)nb + c.g.p->root_b0);
    const int bbeg = root ? (int)min<long long>(rb0, (long long)nb) : 0;

    // remove one vertex, main.cu:1408-1462; returns true on a member failure
    // With members_only, candidates keep their exdeg and `pend` names the
    // removed vertex: the child build applies it to them in its own pass.
    int pend = -1;
    auto remove_step = [&](bool count, bool members_only) -> bool {
        int mdg = get_mindeg(c, nm);
        nc--;
        tp--;
        int rr = PM(tp);
        pend = members_only ? rr : -1;
        bool fail = false;
        int fid = INT32_MAX;  // smallest failing member id
        const int upto = members_only ? nm : tp;
        for (int i = c.lane; i < upto; i += WAVE) {
            if (gbit(c.M1, c.W, rr, PM(i))) {
                int ex = w.pex[i] - 1;
                w.pex[i] = ex;
                if (i < nm && w.pin[i] + ex < mdg) {
                    fail = true;
                    fid = min(fid, w.pid[i]);
                }
            }
        }
        wsync();
        if (ballot(fail)) {
            if (count) {
                // the reference scans N(r) in ascending id order and stops
                // after the first failing member
                fid = uni(wave_min(fid));
                const int rid = w.pid[tp];
                const uint64_t* row = c.g.p->A1 + (size_t)rid * c.g.p->rw;
                int rank = 0;
                for (int wd = c.lane; wd <= (fid >> 6); wd += WAVE) {
                    uint64_t x = row[wd];
                    if (wd == (fid >> 6)) x &= (1ull << (fid & 63)) - 1ull;
                    rank += __popcll(x);
                }
                wadj_lanes(c, rank + (c.lane == 0));
            }
            return true;
        }
        if (count) wadj_add(c, c.g.p->d1r[w.pid[tp]]);
        return false;
    };

    // exdeg of positions [0, keep) minus their neighbours among positions
    // [keep, keep + nr): several removals at once
    auto bulk_remove = [&](int keep, int nr) {
        if (nr <= 0) return;
        if (nr < c.W) {
            for (int k = c.lane; k < nr; k += WAVE) w.adjc[k] = PM(keep + k);
            wsync();
            for (int i = c.lane; i < keep; i += WAVE) {
                const int row = PM(i);
                int cnt = 0;
                for (int k = 0; k < nr; k++) cnt += (int)gbit(c.M1, c.W, row, w.adjc[k]);
                w.pex[i] -= cnt;
            }
        } else {
            clear_mask(c, w.mk0);
            for (int q = keep + c.lane; q < keep + nr; q += WAVE) set_mask_bit(w.mk0, PM(q));
            wsync();
            for (int i = c.lane; i < keep; i += WAVE) w.pex[i] -= rowpop4(c, c.M1, PM(i), w.mk0);
        }
        wsync();
    };

    int b_lo = 0, b_hi = nb;
    bool count_removes = true;
    // a root's branches are split finer, to about one item per resident wave
    // over all roots: until the roots' items produce children nothing else
    // can run
    int ch = a.split_chunk;
    if (ref.depth == 0 && ch > 0 && a.n_roots > 0) {
        const long long want = ((long long)(bend - bbeg) * a.n_roots + gridDim.x - 1) / gridDim.x;
        ch = (int)max(1ll, min((long long)ch, want));
    }
    if (item) {
        // split item: apply removals 1..b0 at once (the prepare step found no
        // failure before its range), then run branches [b0, b1); b0 and b1
        // count from the record the prepare step stored
        const int b0 = (int)ref.b0;
        if (b0 > 0) {
            bulk_remove(t0 - b0, b0);
            tp = t0 - b0;
            nc -= b0;
        }
        ph_mark(c, PH_REMOVE);
        b_lo = b0;
        b_hi = (int)ref.b1;
        count_removes = false;
    } else if (a.queue_mode && ch > 0 && (uint32_t)t0 > a.split_min_t && bend - bbeg > ch) {
        // prepare step: find the first failing removal, count the
        // removals' scans, store the record as it stands before the first
        // branch of the window (with lvl2adj) and push the items
        // Removal k (branch b = k) takes position t0 - k.  Member i fails at
        // the removal of its need_i-th neighbour counted from the end, with
        // need_i = max(pin + pex - mindeg(nm) + 1, 1); kb = the first such
        // removal over all members (bend: none before the window's end).
        const int mdg = get_mindeg(c, nm);
        int kb = bend;
        for (int i = 0; i < nm; i++) {
            const int need = max(w.pin[i] + w.pex[i] - mdg + 1, 1);
            const int row = PM(i);
            // removals up to kb inclusive: a member failing at the same
            // removal as an earlier one must be seen (smallest failing id)
            const int lim = min(kb, bend - 1);
            int cnt = 0, ki = INT32_MAX;
            for (int k0 = 1; k0 <= lim && ki == INT32_MAX; k0 += WAVE) {
                const int k = k0 + c.lane;
                const bool nb_ = k <= lim && gbit(c.M1, c.W, row, PM(t0 - k));
                const uint64_t bm = ballot(nb_);
                const int pc = __popcll(bm);
                if (cnt + pc >= need) {
                    const uint64_t hit = ballot(nb_ && cnt + lanes_below(bm) + 1 == need);
                    ki = k0 + __builtin_ctzll(hit);
                }
                cnt += pc;
            }
            w.rem[i] = ki;
            kb = min(kb, ki);
        }
        wsync();
        // W_adj of the removals: full scans of N(r) for removals 1..kb-1,
        // the partial scan up to the smallest failing member for removal kb
        long long sw = 0;
        for (int k = 1 + c.lane; k < kb; k += WAVE) sw += c.g.p->d1r[w.pid[t0 - k]];
        wadj_lanes(c, sw);
        if (kb < bend) {
            int fid = INT32_MAX;
            for (int i = c.lane; i < nm; i += WAVE)
                if (w.rem[i] == kb) fid = min(fid, w.pid[i]);
            fid = uni(wave_min(fid));
            const uint64_t* rrow = c.g.p->A1 + (size_t)w.pid[t0 - kb] * c.g.p->rw;
            int rank = 0;
            for (int wd = c.lane; wd <= (fid >> 6); wd += WAVE) {
                uint64_t x = rrow[wd];
                if (wd == (fid >> 6)) x &= (1ull << (fid & 63)) - 1ull;
                rank += __popcll(x);
            }
            wadj_lanes(c, rank + (c.lane == 0));
        }
        const int hi = min(kb, bend);
        if (hi <= bbeg) {
            ph_mark(c, PH_PREPARE);
            return;
        }
        // removals 1..bbeg happen before the window: apply them once here
        const int pre = bbeg;
        const int ts = t0 - pre;
        bulk_remove(ts, pre);
        // the record plus one Vrec after it whose id counts the items still
        // to run (the last one frees the block)
        unsigned long long slot = 0;
        if (c.lane == 0) slot = pool_alloc(c, (uint32_t)ts + 1u);
        slot = rdlane64(slot, 0);
        if (slot == ~0ull) {
            if (c.lane == 0) atomicOr(a.error, 1);
            return;
        }
        for (int i = c.lane; i < ts; i += WAVE) {
            Vrec v;
            v.id = w.pid[i];
            v.label = w.plab[i] | (w.pl2[i] << 4);
            v.indeg = w.pin[i];
            v.exdeg = w.pex[i];
            a.pool[slot + i] = v;
        }
        const int n_items = (hi - bbeg + ch - 1) / ch;
        if (c.lane == 0) a.pool[slot + ts] = Vrec{n_items, 0, 0, 0};
        // item boundaries balance the work rather than the branch count:
        // branch b builds a child from about t0 - b vertices, so with
        // S(b) = sum of (t0 - x) over x in [bbeg, b), boundary j is the first
        // b with S(b) >= j S(hi) / n_items (kept strictly increasing)
        const long long t0l = t0;
        auto S = [&](long long b) {
            return (b - bbeg) * t0l - (b * (b - 1) - (long long)bbeg * (bbeg - 1)) / 2;
        };
        const long long s_tot = S(hi);
        auto boundary = [&](int j) -> int {
            if (j <= 0) return bbeg;
            if (j >= n_items) return hi;
            const long long target = s_tot * j / n_items;
            int lo = bbeg, up = hi;
            while (lo < up) {
                const int mid = (lo + up) >> 1;
                if (S(mid) >= target)
                    up = mid;
                else
                    lo = mid + 1;
            }
            return min(max(lo, bbeg + j), hi - (n_items - j));
        };
        auto item_ref = [&](int j) {
            TaskRef r;
            r.off = slot;
            r.t = (uint32_t)ts;
            r.depth = ref.depth;
            r.inst = ref.inst;
            r.bound = (uint64_t)c.slot;  // the task the item belongs to (trace)
            r.b0 = (uint32_t)(boundary(j) - pre);
            r.b1 = (uint32_t)(boundary(j + 1) - pre);
            return r;
        };
        // the items go to the top ring in chunks of up to 64 claimed at once
        // (lanes write the refs, one release fence per chunk); what the ring
        // cannot take stays on this wave's stack
        const int top = a.n_buckets - 1;
        const int rg = top * a.q_shards + c.shard;
        int done = 0;
        while (done < n_items) {
            unsigned long long pos = 0;
            const int k = ring_claim(c, rg, min(n_items - done, WAVE), &pos);
            if (k == 0) break;
            if (c.lane == 0)
                __hip_atomic_fetch_add(a.pending, (unsigned long long)k, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            if (c.lane < k) {
                const unsigned long long gi = (unsigned long long)rg * a.q_cap + (pos + c.lane) % a.q_cap;
                if (a.trace_parent) a.trace_parent[gi] = -2;
                a.q_refs[gi] = item_ref(done + c.lane);
            }
            ring_publish(c, rg, pos, k);
            done += k;
        }
        for (; done < n_items; done++)
            if (!push_ref(c, item_ref(done), top, false, nullptr)) break;
        ph_mark(c, PH_PREPARE);
        return;
    }

    // A window that starts at branch bbeg > 0 first needs removals
    // 1..bbeg.  Without members none of them can fail, so they are applied at
    // once (their adjacency scans counted as the sequential removals count
    // them) and the loop starts at bbeg.
    if (!item && bbeg > 0 && nm == 0 && bbeg < bend) {
        long long sw = 0;
        for (int k = 1 + c.lane; k <= bbeg; k += WAVE) sw += c.g.p->d1r[w.pid[t0 - k]];
        if (count_removes) wadj_lanes(c, sw);
        bulk_remove(t0 - bbeg, bbeg);
        tp = t0 - bbeg;
        nc -= bbeg;
        b_lo = bbeg;
        ph_mark(c, PH_REMOVE);
    }

    for (int b = b_lo; b < b_hi; b++) {
        if (!item && b >= bend) break;
        pend = -1;
        if (b > b_lo || (!item && b > 0 && b_lo == 0)) {
            // before the window (no child built) the removal is applied
            // whole; otherwise the child build applies it to the candidates
            const bool stop = remove_step(count_removes, item || b >= bbeg);
            ph_mark(c, PH_REMOVE);
            if (stop) break;
        }
        if (!item && b < bbeg) continue;
        // child record (main.cu:1094-1180 with add_one_vertex, main.cu:1465-1523,
        // and diameter pruning, main.cu:1709-1732, applied while it is built):
        // members, then the branch vertex (parent position tp - 1) as a
        // member, then the candidates inside N2(v) in parent order with their
        // degrees updated for v; candidates outside N2(v) follow with id and
        // matrix index only (label -1: degree pruning counts them as removed)
        if (c.lane == 0) c.st->branches++;
        const int vsrc = tp - 1;
        const int vpp = PM(vsrc);
        wadj_add(c, (long long)c.g.p->d1r[w.pid[vsrc]] + c.g.p->d2r[w.pid[vsrc]]);
        for (int i = c.lane; i < nm; i += WAVE) {
            const int pp = PM(i);
            const int a1 = (int)gbit(c.M1, c.W, vpp, pp);
            w.cid[i] = w.pid[i];
            w.cpp[i] = pp;
            w.clab[i] = w.plab[i];
            w.cin[i] = w.pin[i] + a1;
            w.cex[i] = w.pex[i] - a1;
            if (c.keys) w.cl2[i] = w.pl2[i];
        }
        if (c.lane == 0) {
            int pv = w.pex[vsrc];
            if (pend >= 0 && gbit(c.M1, c.W, pend, vpp)) w.pex[vsrc] = --pv;
            w.cid[nm] = w.pid[vsrc];
            w.cpp[nm] = vpp;
            w.clab[nm] = 1;
            w.cin[nm] = w.pin[vsrc];
            w.cex[nm] = pv;
            if (c.keys) w.cl2[nm] = w.pl2[vsrc];
        }
        int m = nm + 1, cc = nc - 1;
        // one pass over the candidates: the survivors of the 2-hop test go
        // to [m, m + ns) in parent order with their degrees updated for v;
        // the others (id and matrix index only) are staged in rem / rmv and
        // appended after the survivors
        int nsd = 0, ndd = 0, hV = 0, hc = 0;
        for (int base = nm; base < vsrc; base += WAVE) {
            const int q = base + c.lane;
            const bool in = q < vsrc;
            const int pp = in ? PM(q) : 0;
            const bool ok = in && gbit(c.M2, c.W, vpp, pp);
            const uint64_t bo = ballot(ok), bx = ballot(in && !ok);
            int pexq = 0;
            if (in) {
                // the removal of this branch (main.cu:1437-1450) for the
                // parent's candidate q
                pexq = w.pex[q];
                if (pend >= 0 && gbit(c.M1, c.W, pend, pp)) w.pex[q] = --pexq;
            }
            int xin = -1;
            if (ok) {
                const int r = lanes_below(bo);
                const int d = m + nsd + r;
                const int a1 = (int)gbit(c.M1, c.W, vpp, pp);
                const int cin_ = w.pin[q] + a1;
                xin = cin_;
                w.cid[d] = w.pid[q];
                w.cpp[d] = pp;
                w.clab[d] = 0;
                w.cin[d] = cin_;
                w.cex[d] = pexq - a1;
                if (c.keys) w.cl2[d] = w.pl2[q];
                if (m >= WAVE) w.cind[nsd + r] = cin_;  // else the register table
            } else if (in) {
                const int d = ndd + lanes_below(bx);
                w.rem[d] = w.pid[q];
                w.rmv[d] = pp;
            }
            nsd += __popcll(bo);
            ndd += __popcll(bx);
            if (m < WAVE) hist_chunk(c, xin, hc, hV);
        }
        const int ns = nsd;
        wsync();
        for (int k = c.lane; k < ndd; k += WAVE) {
            const int d = m + ns + k;
            w.cid[d] = w.rem[k];
            w.cpp[d] = w.rmv[k];
            w.clab[d] = -1;
        }
        int tc = m + ns + ndd;
        int nind = ns;
        wsync();
        ph_mark(c, PH_ADD);
        int ub = 0, lb = 0, med = 0;
        // degree pruning, then critical-vertex pruning, which ends in a
        // second degree pruning (one call site: the call is inlined)
        bool first = true, map_empty = false, skip = false;
        int r = 0;
        int preV = m < WAVE ? hV : -1;  // the build's table of the survivors' indegrees
        for (;;) {
            const bool fail = degree_pruning(c, tc, cc, m, ub, lb, med, nind, map_empty, preV, hc);
            preV = -1;
            if (!first) {
                r = fail ? 1 : 0;
                break;
            }
            if (fail) {
                if (m >= c.g.min_size) check_for_clique(c, m);
                skip = true;
                break;
            }
            const int crit = critical_prep(c, tc, cc, m, lb, nind, map_empty);
            ph_mark(c, PH_CRIT);
            if (crit == 2) {
                skip = true;
                break;
            }
            first = false;
        }
        if (skip) {
            ph_mark(c, PH_PRUNE);
            continue;
        }
        if (m >= c.g.min_size) check_for_clique(c, m);
        ph_mark(c, PH_PRUNE);
        if (r == 1) continue;
        if (cc > 0) write_child(c, tc, m, ub - lb);
        wsync();
        ph_mark(c, PH_CHILD);
    }
#undef PM
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ void flush_stats(Ctx& c, DevStats* d) {
    const unsigned long long wadj = (unsigned long long)wave_sum64(c.wadj_l);
    c.wadj_l = 0;
    if (c.lane == 0) {
        const LStats& t = *c.st;
        if (t.tasks) atomicAdd(&d->tasks, t.tasks);
        if (t.branches) atomicAdd(&d->branches, t.branches);
        if (t.children) atomicAdd(&d->children, t.children);
        if (t.raw) atomicAdd(&d->raw_cliques, t.raw);
        if (t.sum_t) atomicAdd(&d->sum_t, t.sum_t);
        if (t.sum_tw) atomicAdd(&d->sum_t_written, t.sum_tw);
        if (t.sum_clq) atomicAdd(&d->sum_clique_len, t.sum_clq);
        if (t.large) atomicAdd(&d->large_tasks, t.large);
        if (t.beyond) atomicAdd(&d->beyond, t.beyond);
        if (wadj) atomicAdd(&d->w_adj, wadj);
        *c.st = LStats{0, 0, 0, 0, 0, 0, 0, 0, 0};
    }
}
__device__ __forceinline__ void flush_stats(Ctx& c) { flush_stats(c, c.a->stats); }

template <bool LOCAL>
__device__ __forceinline__ void task_loop(Ctx& c) {
    const ExpandArgs& a = *c.a;
    const uint32_t nw = gridDim.x;
    uint32_t i;
    if (a.static_sched) {
        i = blockIdx.x;
    } else {
        uint32_t x = 0;
        if (c.lane == 0) x = atomicAdd(a.task_counter, 1u);
        i = (uint32_t)rdlane((int)x, 0);
    }
    while (i < a.n_in) {
        TaskRef ref = a.in_refs[i];
        if (ref.t <= c.cap)
            process_task<LOCAL>(c, ref);
        else if (c.lane == 0)
            atomicOr(a.error, 2);
        wsync();
        if (a.static_sched) {
            i += nw;
        } else {
            uint32_t x = 0;
            if (c.lane == 0) x = atomicAdd(a.task_counter, 1u);
            i = (uint32_t)rdlane((int)x, 0);
        }
    }
    flush_stats(c);
}

template <int TL>
__global__ __launch_bounds__(64) void expand_local(const ExpandArgs* __restrict__ ap) {
    const ExpandArgs& a = *ap;
    constexpr size_t BYTES = ws_bytes(TL, true, (TL + 63) / 64);
    __shared__ __attribute__((aligned(16))) char lds[BYTES];
    Ctx c;
    c.wadj_l = 0;
    c.a = ap;
    c.g = load_graph(&ap->g);
    c.phase_on = ap->phase != nullptr;
    c.inst = 0;
    c.slot = -1;
    c.lane = lane_id();
    c.w = make_ws(lds, TL, true, (TL + 63) / 64);
    c.cap = TL;
    __shared__ LStats s_st;
    __shared__ unsigned long long s_ph[PH_COUNT];
    c.st = &s_st;
    c.ph = s_ph;
    if (c.lane == 0) s_st = LStats{0, 0, 0, 0, 0, 0, 0, 0, 0};
    ph_start(c);
    task_loop<true>(c);
}

__global__ __launch_bounds__(64) void expand_hbm(const ExpandArgs* __restrict__ ap) {
    const ExpandArgs& a = *ap;
    Ctx c;
    c.wadj_l = 0;
    c.a = ap;
    c.g = load_graph(&ap->g);
    c.phase_on = ap->phase != nullptr;
    c.inst = 0;
    c.slot = -1;
    c.lane = lane_id();
    char* base = a.slab + (size_t)blockIdx.x * a.slab_bytes;
    c.w = make_ws(base, a.slab_tcap, false, a.slab_mwords);
    c.cap = a.slab_tcap;
    __shared__ LStats s_st;
    __shared__ unsigned long long s_ph[PH_COUNT];
    c.st = &s_st;
    c.ph = s_ph;
    if (c.lane == 0) s_st = LStats{0, 0, 0, 0, 0, 0, 0, 0, 0};
    ph_start(c);
    task_loop<false>(c);
}

// ---------------------------------------------------------------------------
// persistent kernel: every wave pops task refs from the FIFO queue until the
// pending counter reaches zero; tasks of at most TL vertices run from LDS,
// larger ones from the wave's HBM slab.
// ---------------------------------------------------------------------------
// (two resident waves per SIMD: the register allocation is held to 256)
template <int TL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void expand_persistent(
    const ExpandArgs* __restrict__ ap) {
    const ExpandArgs& a = *ap;
    // dynamic LDS: max(TL workspace with local matrices, arrays of wide tasks)
    extern __shared__ __attribute__((aligned(16))) char lds[];
    Ctx c;
    c.wadj_l = 0;
    c.a = ap;
    c.g = load_graph(&ap->g);
    c.phase_on = ap->phase != nullptr;
    c.inst = 0;
    c.slot = -1;
    c.lane = lane_id();
    c.pstack = a.priv + (size_t)blockIdx.x * a.priv_cap;
    c.pbkt = a.priv_bkt + (size_t)blockIdx.x * a.priv_cap;
    c.pbot = 0;
    c.ptop = 0;
    c.shard = (int)(blockIdx.x % (unsigned)a.q_shards);
    // this wave's stacks of free blocks of the small classes
    __shared__ uint32_t s_fl[FL_SMALL * FL_LOCAL_MAX];
    __shared__ uint32_t s_fln[FL_SMALL];
    c.fl = s_fl;
    c.fln = s_fln;
    if (c.lane < FL_SMALL) s_fln[c.lane] = 0;
    wsync();
    char* const slab = a.slab + (size_t)blockIdx.x * a.slab_bytes;
    const uint32_t mid = a.mid_tcap;
    const uint32_t wide = a.lds_wide;
    __shared__ LStats s_st;
    __shared__ unsigned long long s_ph[PH_COUNT];
    c.st = &s_st;
    c.ph = s_ph;
    if (c.lane == 0) s_st = LStats{0, 0, 0, 0, 0, 0, 0, 0, 0};
    ph_start(c);
    // consecutive idle polls before the launch is declared stuck (about a
    // minute of back-off): every wave then leaves with error bit 4
    constexpr int IDLE_LIMIT = 1 << 22;
    const unsigned long long G = gridDim.x;
    int idle = 0;       // consecutive empty polls (back-off)
    int idle_total = 0;
    int sinst = -1;     // instance the wave's counters belong to
    // static order: this wave takes positions blockIdx.x, +G, +2G, ... of
    // the single ring
    unsigned long long spos = blockIdx.x;

    // positions claimed from a ring and not yet run: their refs wait in LDS
    // (registers are the scarce resource of this kernel)
    __shared__ TaskRef s_claim[4];
    __shared__ TaskRef s_top;
    c.ptopref = &s_top;
    c.ptopi = ~0u;
    c.share_v = 0;  // reads as nobody waiting
    c.share_n = 0;
    c.share_left = 0;
    c.share_ok = false;
    int nq = 0, qi = 0;
    unsigned long long qbase = 0, qh = 0;
    // `pending` counts the tasks waiting in the rings plus the waves that hold
    // work (claimed positions or a non-empty stack); `hungry` counts the
    // waves that hold none.  A wave changes them only when it takes work
    // from the rings, hands work to them or runs out, not per task.
    bool busy = false;
    for (;;) {
        TaskRef ref;
        long long gslot;
        if (qi < nq) {
            ref = s_claim[qi];
            if (c.phase_on && c.lane == 0) c.ph[PH_N_RING]++;
            gslot = (long long)(qbase + (qh + (unsigned long long)qi) % a.q_cap);
            qi++;
        } else if (a.static_sched && c.ptop != c.pbot && static_flush(c)) {
            continue;
        } else if (c.ptop != c.pbot) {
            // this wave's own stack (depth first)
            idle = 0;
            ph_mark(c, PH_POP);
            if (!a.static_sched && c.ptop - c.pbot > 1) {
                // more waves wait for work than the rings hold: hand over
                // the oldest (shallowest, so largest) records, up to 16 at
                // once, into the bucket of the oldest one.  The rings hold
                // pending - (G - hungry) tasks; handing over more than the
                // waiting waves take turns the search breadth first and
                // multiplies the live records.
                unsigned long long ph = 0;
                if (c.lane == 0) ph = __hip_atomic_load(a.pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ph = rdlane64(ph, 0);
                const long long hg = (long long)(ph >> 32), pd = (long long)(ph & 0xFFFFFFFFull);
                const long long waiting = pd - ((long long)G - hg);
                const long long need = hg - (waiting > 0 ? waiting : 0);
                if (need > 0) {
                    const int want = (int)min<long long>(min<long long>((long long)(c.ptop - c.pbot - 1), 16ll), need);
                    if (a.n_buckets > 1) {
                        // cost order: the 16 oldest records are ordered by
                        // predicted bucket, largest first (stable), so the
                        // hand-over takes the ones the model predicts largest.
                        // Lane j holds record j; its rank is the number of
                        // records before it in that order.
                        wsync();  // lane 0's stack stores, visible to every lane
                        const int span = (int)min<uint32_t>(c.ptop - c.pbot - 1, 16u);
                        const uint32_t at = (c.pbot + (uint32_t)c.lane) % a.priv_cap;
                        const int bj = c.lane < span ? (int)c.pbkt[at] : -1;
                        int rank = 0;
                        for (int i = 0; i < span; i++) {
                            const int bi = rdlane(bj, i);
                            rank += (bi > bj || (bi == bj && i < c.lane)) ? 1 : 0;
                        }
                        if (ballot(c.lane < span && rank != c.lane)) {
                            TaskRef tr;
                            if (c.lane < span) tr = c.pstack[at];
                            wsync();
                            if (c.lane < span) {
                                const uint32_t to = (c.pbot + (uint32_t)rank) % a.priv_cap;
                                c.pstack[to] = tr;
                                c.pbkt[to] = (unsigned char)bj;
                            }
                            wsync();
                        }
                    }
                    int bk = 0;
                    if (c.lane == 0) bk = c.pbkt[c.pbot % a.priv_cap];
                    const int rg = rdlane(bk, 0) * a.q_shards + c.shard;
                    // a hand-over is optional: one attempt, and when another
                    // producer moved the tail first the wave keeps its records
                    // (retrying under contention cost more than the records
                    // were worth on trees of tiny tasks)
                    unsigned long long pos = 0;
                    const int k = ring_claim(c, rg, want, &pos, 2);
                    if (k > 0) {
                        // lane j copies entry j (the barrier makes lane 0's
                        // stack stores visible to the other lanes)
                        wsync();
                        if (c.lane < k)
                            a.q_refs[(unsigned long long)rg * a.q_cap + (pos + c.lane) % a.q_cap] =
                                c.pstack[(c.pbot + (uint32_t)c.lane) % a.priv_cap];
                        if (c.lane == 0)
                            __hip_atomic_fetch_add(a.pending, (unsigned long long)k,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ring_publish(c, rg, pos, k);
                        c.pbot += (uint32_t)k;
                        if (c.phase_on && c.lane == 0) c.ph[PH_N_GIVEN] += (unsigned long long)k;
                    }
                }
            }
            ph_mark(c, PH_GIVE);
            c.ptop--;
            TaskRef r;
            // the newest entry is usually still in LDS (pushed last)
            if (c.lane == 0) r = c.ptopi == c.ptop ? *c.ptopref : c.pstack[c.ptop % a.priv_cap];
            c.ptopi = ~0u;
            ref = bcast_ref(r, 0);
            gslot = -1;
            if (c.phase_on && c.lane == 0) c.ph[PH_N_OWN]++;
        } else {
            // ---- the shared rings ----
            int state = 0;  // 0 nothing published, 1 got positions, 2 done, 3 retry now
            unsigned long long h = 0;
            int nk = 1;     // consecutive positions claimed
            int bsel = 0;   // their ring
            if (a.static_sched) {
                if (c.lane == 0) {
                    const unsigned long long sq =
                        __hip_atomic_load(&a.q_seq[spos % a.q_cap], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (sq == spos + 1ull) {
                        state = 1;
                        h = spos;
                        spos += G;
                    }
                }
            } else {
                // every ring at once: lane l reads the head and tail of ring l
                // (bucket l / q_shards, shard l % q_shards)
                const int NS = a.q_shards;
                const int NR = a.n_buckets * NS;
                unsigned long long hl = 0, tl = 0;
                if (c.lane < NR) {
                    hl = __hip_atomic_load(&a.q_head[c.lane * QSTRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    tl = __hip_atomic_load(&a.q_tail[c.lane * QSTRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                const uint64_t m = ballot(c.lane < NR && tl > hl);
                if (m) {
                    // this wave's own shard first (its XCD's L2 holds the
                    // records its neighbours wrote), the highest bucket with
                    // work there; else the highest bucket with work anywhere,
                    // from the next shard after this one
                    uint64_t own = 0;
                    for (int q = 0; q < a.n_buckets; q++) own |= 1ull << (q * NS + c.shard);
                    const uint64_t mo = m & own;
                    const int b = (63 - __builtin_clzll(mo ? mo : m)) / NS;
                    const uint64_t full = NS == 64 ? ~0ull : (1ull << NS) - 1;
                    const uint64_t sm = (m >> (b * NS)) & full;
                    const uint64_t rot = ((sm >> c.shard) | (sm << ((NS - c.shard) & 63))) & full;
                    const int rsel = b * NS + (c.shard + __builtin_ctzll(rot)) % NS;
                    const unsigned long long hb = rdlane64(hl, rsel), tb = rdlane64(tl, rsel);
                    if (c.lane == 0) {
                        // a very deep ring hands out several consecutive
                        // positions per claim: one head update per claim limits
                        // how fast tiny tasks can be taken.  Only very deep: the
                        // positions of one claim run one after another, which
                        // lengthens the critical path when they are heavy.  Only
                        // published positions are claimed.
                        const unsigned long long avail = tb - hb;
                        const unsigned long long per = G / (unsigned long long)NS + 1;
                        const int k = avail >= 32 * per ? 4 : avail >= 16 * per ? 2 : 1;
                        const unsigned long long* seq = a.q_seq + (size_t)rsel * a.q_cap;
                        int pub = 0;
                        while (pub < k && __hip_atomic_load(&seq[(hb + pub) % a.q_cap], __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT) == hb + pub + 1ull)
                            pub++;
                        if (pub == 0) {
                            state = 3;  // claimed by a producer, about to be published
                        } else {
                            unsigned long long e = hb;
                            if (__hip_atomic_compare_exchange_strong(&a.q_head[rsel * QSTRIDE], &e,
                                                                     hb + (unsigned long long)pub, __ATOMIC_RELAXED,
                                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                                state = 1;
                                h = hb;
                                nk = pub;
                                bsel = rsel;
                            } else {
                                state = 3;
                            }
                        }
                    }
                }
            }
            // every ring looked empty: finished once nothing is pending (no
            // ring task and no busy wave: a wave counts itself out only when
            // its stack is empty, after pushing what it made).  Polls are
            // relaxed: an acquire invalidates the CU's vector L1 under the
            // waves still working.
            if (c.lane == 0 && state == 0 &&
                (__hip_atomic_load(a.pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0xFFFFFFFFull) == 0)
                state = 2;
            state = rdlane(state, 0);
            if (state == 2) break;
            if (state == 3) continue;
            if (state == 0) {
                if (busy) {
                    // out of work: no longer counted in pending, now hungry
                    busy = false;
                    if (c.lane == 0) {
                        __hip_atomic_fetch_add(a.pending, PH_HUNGRY - 1ull, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    }
                    continue;  // look once more before sleeping
                }
                // a lost task or another wave's fault: leave (error bit 4)
                if (rdlane(c.lane == 0 ? __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0,
                           0) & 4)
                    break;
                idle = idle < 16 ? idle + 1 : 16;
                if (++idle_total >= IDLE_LIMIT) {
                    if (c.lane == 0) atomicOr(a.error, 4);
                    break;
                }
                for (int k = 0; k < idle; k++) __builtin_amdgcn_s_sleep(32);
                ph_mark(c, PH_IDLE);
                continue;
            }
            idle = 0;
            idle_total = 0;
            h = rdlane64(h, 0);
            nk = rdlane(nk, 0);
            bsel = rdlane(bsel, 0);
            // the claimed tasks leave the rings; an idle wave becomes busy
            if (c.lane == 0) {
                const unsigned long long d =
                    (unsigned long long)((busy ? 0 : 1) - (long long)nk) - (busy ? 0ull : PH_HUNGRY);
                if (d) __hip_atomic_fetch_add(a.pending, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            busy = true;
            // one acquire per claim (pairs with the producers' release before
            // the sequence words), then lane q reads the ref of position h + q
            // and gives the cell back for the next lap
            ph_mark(c, PH_POP);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const unsigned long long base = (unsigned long long)bsel * a.q_cap;
            if (c.lane < nk) s_claim[c.lane] = a.q_refs[base + (h + c.lane) % a.q_cap];
            wsync();
            // the refs are in LDS (the barrier waited for their loads), so
            // the cells go back with plain relaxed stores: no release, whose
            // agent-scope form writes back this XCD's L2
            if (c.lane < nk)
                __hip_atomic_store(&a.q_seq[base + (h + c.lane) % a.q_cap], h + (unsigned long long)c.lane + a.q_cap,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ph_mark(c, PH_SYNC);
            nq = nk;
            qi = 0;
            qbase = base;
            qh = h;
            continue;
        }
        ph_mark(c, PH_POP);
        const long long tl_start = a.timeline ? wall_clock64() : 0;
        if (a.graphs) {
            c.g = load_graph(&a.graphs[ref.inst]);
            c.inst = ref.inst;
        }
        if (a.inst_stats && (int)ref.inst != sinst) {
            if (sinst >= 0) flush_stats(c, &a.inst_stats[sinst]);
            sinst = ref.inst;
        }
        c.slot = ref.b1 > 0 ? (long long)ref.bound : gslot;
        c.share_ok = ref.t >= SHARE_MIN_T;
        // workspace: LDS with local matrices, LDS arrays with root
        // matrices, slab with local matrices, or slab in root space
        const bool fits = ref.t <= a.slab_tcap;
        const bool local = ref.t <= (uint32_t)TL || (ref.t > wide && ref.t <= mid);
        // (built per task: only one workspace descriptor is live)
        if (ref.t <= (uint32_t)TL) {
            c.w = make_ws(lds, TL, true, (TL + 63) / 64);
            c.cap = TL;
        } else if (ref.t <= wide) {
            c.w = make_ws_split(lds, wide,
                                a.lds_masks >= 0 ? (uint64_t*)(lds + a.lds_masks)
                                                 : (uint64_t*)(slab + ws_bytes(a.slab_tcap, false, 0)),
                                a.slab_mwords);
            c.cap = wide;
        } else if (ref.t <= mid) {
            c.w = make_ws(slab, mid, true, (mid + 63) / 64);
            if (a.lds_masks >= 0) {
                c.w.mk0 = (uint64_t*)(lds + a.lds_masks);
                c.w.mk1 = c.w.mk0 + (mid + 63) / 64;
            }
            c.cap = mid;
        } else {
            c.w = make_ws(slab, a.slab_tcap, false, a.slab_mwords);
            if (a.lds_masks >= 0) {
                c.w.mk0 = (uint64_t*)(lds + a.lds_masks);
                c.w.mk1 = c.w.mk0 + a.slab_mwords;
                // the LDS below the masks (the wide tier's arrays) is free
                // during a slab task: as many of the task's arrays as fit go
                // there, in slots of S = t rounded up to 64 entries plus 64
                // (every array and list of a task holds at most t + 2
                // entries: positions below t, and the LU tables over
                // indegrees 0..V+1 with V <= members), in this order: the
                // child's arrays every pruning pass reads (indegree,
                // exdegree, label, id, matrix index, lvl2adj), the pruning
                // lists (rem, rmv, cind, adjc), then the parent's arrays
                // (indegree, exdegree, label, id, lvl2adj).  When the
                // children lie past the depth window their lvl2adj is never
                // kept, and the parent's exdegree, updated by every branch,
                // takes its slot.  The sort keys keep their slab addresses
                // (the bitonic sort pads them to a power of two).
                const uint32_t S = ((ref.t + 63u) & ~63u) + 64u;
                const uint32_t k = (uint32_t)a.lds_masks / (S * 4);
                const bool keys = !(c.g.max_depth >= 0 && (int)ref.depth + 1 > c.g.max_depth);
                int* lp = (int*)lds;
                uint32_t n = 0;
                auto put = [&](int*& p) {
                    if (n < k) p = lp + n * S;
                    n++;
                };
                put(c.w.cin);
                put(c.w.cex);
                put(c.w.clab);
                put(c.w.cid);
                put(c.w.cpp);
                put(keys ? c.w.cl2 : c.w.pex);
                put(c.w.rem);
                put(c.w.rmv);
                put(c.w.cind);
                put(c.w.adjc);
                if (k > 8) c.w.T2 = (int)S;  // the LU tables' capacity (cind, adjc)
                put(c.w.pin);
                if (keys) put(c.w.pex);
                put(c.w.plab);
                put(c.w.pid);
                if (keys) put(c.w.pl2);
            }
            c.cap = a.slab_tcap;
        }
        if (!fits) {
            if (c.lane == 0) atomicOr(a.error, 2);
        } else if (local) {
            process_task<true>(c, ref);
        } else {
            process_task<false>(c, ref);
        }
        wsync();
        if (a.timeline && gslot >= 0 && c.lane == 0) {
            unsigned long long* r = a.timeline + 4 * (unsigned long long)gslot;
            r[0] = (unsigned long long)tl_start;
            r[1] = (unsigned long long)wall_clock64();
            r[2] = ((unsigned long long)blockIdx.x << 32) | ref.t;
            r[3] = (unsigned long long)ref.depth | ((unsigned long long)ref.b0 << 16) |
                   ((unsigned long long)ref.b1 << 40);
        }
        ph_mark(c, PH_CHILD);
        if (c.lane == 0) {
            // the record's block: a split item's prepared record is freed by
            // the last of its items (count in the Vrec after the record);
            // roots live outside the pool's blocks
            if (ref.b1 > 0) {
                const int left = __hip_atomic_fetch_add(&a.pool[ref.off + ref.t].id, -1, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                if (left == 1) pool_free(c, ref.off, ref.t + 1u);
            } else if (ref.depth > 0) {
                pool_free(c, ref.off, ref.t);
            }
        }
        ph_mark(c, PH_POOL);
    }
    ph_mark(c, PH_IDLE);
    ph_flush(c);
    if (sinst >= 0)
        flush_stats(c, &a.inst_stats[sinst]);
    else
        flush_stats(c);
}

}  // namespace qcdev

template __global__ void qcdev::expand_persistent<64>(const ExpandArgs*);
template __global__ void qcdev::expand_persistent<128>(const ExpandArgs*);
template __global__ void qcdev::expand_persistent<256>(const ExpandArgs*);

// explicit instantiations exported for the host driver
template __global__ void qcdev::expand_local<64>(const ExpandArgs*);
template __global__ void qcdev::expand_local<128>(const ExpandArgs*);
template __global__ void qcdev::expand_local<256>(const ExpandArgs*);
