// qc_oracle.cpp -- CPU restatement of cuQC's host expand/prune path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the qcmine HIP
// path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it (liboracle.so), and only to compare against or to time.  Nothing in
// cuqc_amd/ links, loads or calls it.
//
// Every function restates one host routine of the reference
// (Mike12041204/cuQC, main.cu); the line ranges are cited per function.
// The restatement follows the *host* column of the reference wherever its
// host and device code diverge.
//
// Map semantics.  The reference keeps one global id->position map
// (vertex_order_map, main.cu:742-750) that every step sets and resets over
// explicit ranges, with a full O(|V|) reset at the end of every
// add_one_vertex (main.cu:1513-1515).  This oracle performs exactly the same
// set/reset ranges, and replaces the O(|V|) reset by resetting every entry
// set since the previous full reset (same resulting state, O(touched) cost).
// Two modes:
//   faithful = 1 : the map state carries over between tasks exactly as in
//                  the reference.  After a successful critical-vertex step
//                  the reference resets only the compacted prefix
//                  (main.cu:1694-1697), so ids removed by that step keep
//                  stale positions until the next full reset.  Inside a task
//                  the next remove_one_vertex overwrites them; after the
//                  root's last branch they reach the next task processed.
//   faithful = 0 : every task starts from a clean map (the stale entries
//                  are cleared when critical_vertex_pruning returns).  This
//                  is what every task of the reference sees except the one
//                  processed right after the root.
// Any stale position that would index outside the current array (heap
// out-of-bounds in the reference) is counted in stats.stale_oob and skipped.
//
// Graph lists.  Expansion only ever looks up N/N2 entries in the map, which
// holds task vertices (a subset of the root task).  The oracle therefore
// restricts N and N2 to root vertices after building the root; N2 keeps the
// full-graph 2-hop semantics of adjToSer.cpp:115-156 (N(v) u N(N(v)) \ {v}).
// W_adj counts entries of these restricted lists.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace {

struct Vtx {
    int id;
    int label;  // 0 cand, 1 member, 2 covered, 3 cover, 4 critical-adjacent, -1 pruned
    int indeg;
    int exdeg;
    int l2;     // lvl2adj
};

struct Stats {
    uint64_t tasks;          // task records expanded (root included)
    uint64_t branches;       // add_one_vertex calls
    uint64_t children;       // child records written
    uint64_t raw_cliques;    // raw sets emitted (lookahead + clique test)
    uint64_t sum_t;          // sum of task sizes over expanded tasks
    uint64_t sum_t_written;  // sum of child sizes written
    uint64_t sum_clique_len; // sum of raw set sizes
    uint64_t w_adj;          // restricted adjacency entries scanned
    uint64_t lookahead_hits; // tasks finished by lookahead
    uint64_t stale_oob;      // faithful mode: stale positions out of range
    uint64_t stale_hits;     // faithful mode: stale positions used
    uint64_t max_t;          // largest task
    uint64_t max_depth;
    uint64_t root_size;
    uint64_t init_w_adj;     // entries scanned while building the root
    uint64_t beyond;         // records past the depth window (not expanded)
    uint64_t root_branches;  // branches of the root task (window range)
    // how often each pruning rule fired (coverage of the rules by a run)
    uint64_t remove_breaks;    // remove_one_vertex stopped at a failing member (main.cu:1437-1450)
    uint64_t lu_low;           // LU bounds with min member indegree below mindeg[m] (main.cu:1880)
    uint64_t lu_lb_invalid;    // no feasible lower bound (main.cu:1895-1900)
    uint64_t lu_sum_invalid;   // the sorted-indegree sum test failed (main.cu:1932-1936)
    uint64_t lu_ub_tightened;  // ub lowered by the sorted-indegree sums (main.cu:1946-1958)
    uint64_t lu_invalid;       // bounds reported invalid
    uint64_t member_fail;      // a member failed vert_isextendable_LU
    uint64_t dp_recompute;     // degree-pruning rounds recomputing exdeg from the remaining set
    uint64_t dp_decrement;     // rounds decrementing exdeg through the removed set
    uint64_t crit_none;        // critical step with no critical-adjacent candidate (empty map, main.cu:1575-1602)
    uint64_t crit_fail;        // critical candidates rejected by the adjacency checks
    uint64_t crit_ok;          // critical candidates added as members
};

struct TraceRow {
    float f[12];
    int64_t parent;
    int64_t beyond;  // children past the depth window (counted, not expanded)
};

// Q order of main.cu:2059-2104: label rank 1 < 2 < 0 < 3 < other, then
// indeg desc, exdeg desc, lvl2adj desc, vertexid desc (a total order).
static inline int label_rank(int l) {
    return l == 1 ? 0 : l == 2 ? 1 : l == 0 ? 2 : l == 3 ? 3 : 4;
}
static inline bool q_less(const Vtx& a, const Vtx& b) {
    int ra = label_rank(a.label), rb = label_rank(b.label);
    if (ra != rb) return ra < rb;
    if (a.indeg != b.indeg) return a.indeg > b.indeg;
    if (a.exdeg != b.exdeg) return a.exdeg > b.exdeg;
    if (a.l2 != b.l2) return a.l2 > b.l2;
    return a.id > b.id;
}

struct Oracle {
    // graph (full)
    int n = 0;
    const uint64_t* off = nullptr;
    const int* adj = nullptr;
    const uint64_t* off2 = nullptr;  // optional full N2 lists
    const int* adj2 = nullptr;
    // restricted lists (root vertices only): storage, and the pointers every
    // step reads (worker copies of a multi-threaded run share the storage of
    // the oracle that built it)
    std::vector<uint64_t> roff_s, r2off_s;
    std::vector<int> radj_s, r2adj_s;
    const uint64_t* roff = nullptr;
    const uint64_t* r2off = nullptr;
    const int* radj = nullptr;
    const int* r2adj = nullptr;
    std::vector<char> in_root;

    double gamma = 0.9;
    int ms = 2;
    std::vector<int> mindeg;
    bool faithful = false;
    bool trace = false;
    // search window: root branches b in [rb0, rb1) (b = j - covered, the
    // reference's loop order; rb0 < 0 starts that many branches before the
    // end) and tasks of depth <= max_depth are expanded
    long rb0 = 0, rb1 = -1;
    // and the branches [lb0, lb1) of every depth-1 task (same conventions)
    long lb0 = 0, lb1 = -1;
    int max_depth = -1;
    int cur_depth = 0;

    std::vector<int> pos;     // vertex_order_map
    std::vector<int> dirty;   // ids set since the last full reset
    std::vector<int> remaining, removed, cind, adjc;

    Stats st{};
    std::vector<uint64_t> clq_off{0};
    std::vector<int> clq_v;
    std::vector<TraceRow> trace_rows;

    // ---- tables (main.cu:526-533) ----
    void build_tables() {
        size_t sz = 2 * (size_t)std::max(n, ms) + 16;
        mindeg.assign(sz, 0);
        for (size_t i = 1; i < sz; i++) mindeg[i] = (int)std::ceil(gamma * (double)((int)i - 1));
    }
    inline int md(long i) const {
        if (i < 0 || (size_t)i >= mindeg.size()) {
            fprintf(stderr, "oracle: mindeg index %ld out of range\n", i);
            abort();
        }
        return mindeg[i];
    }
    // main.cu:2144-2151
    inline int get_mindeg(long s) const { return s < ms ? md(ms) : md(s); }

    // main.cu:2153-2173
    inline bool cand_valid(const Vtx& v, int m, int ub, int lb, int med) const {
        int tot = v.indeg + v.exdeg;
        if (tot < md(ms)) return false;
        if (tot < get_mindeg((long)m + v.exdeg + 1)) return false;
        if (tot < med) return false;
        if (v.indeg + ub - 1 < md((long)m + lb)) return false;
        if (tot < get_mindeg((long)m + lb)) return false;
        return true;
    }
    // main.cu:2175-2198
    inline bool vert_ext(const Vtx& v, int m, int ub, int lb, int med) const {
        int tot = v.indeg + v.exdeg;
        if (tot < md(ms)) return false;
        if (tot < get_mindeg((long)m + v.exdeg)) return false;
        if (tot < med) return false;
        if (v.exdeg == 0 && v.indeg < get_mindeg((long)m + v.exdeg)) return false;
        if (v.indeg + ub < md((long)m + ub)) return false;
        if (tot < get_mindeg((long)m + lb)) return false;
        return true;
    }

    // ---- map primitives ----
    inline void mset(int id, int p) { pos[id] = p; dirty.push_back(id); }
    inline void munset(int id) { pos[id] = -1; }
    void full_reset() {
        for (int id : dirty) pos[id] = -1;
        dirty.clear();
    }
    // position lookup with stale-range guard; returns -1 if out of range
    inline int look(int id, int t) {
        int p = pos[id];
        if (p >= t) {
            st.stale_oob++;
            return -1;
        }
        return p;
    }

    // ---- LU bounds, main.cu:1857-1972 ----
    bool lu_bounds(int& ub, int& lb, int& med, const Vtx* vs, int m, int nc, const int* ci) {
        bool invalid = false;
        int min_in = vs[0].indeg, min_in_ex = vs[0].exdeg;
        int min_tot = vs[0].indeg + vs[0].exdeg;
        long sum_in = vs[0].indeg;
        for (int i = 1; i < m; i++) {
            sum_in += vs[i].indeg;
            if (vs[i].indeg < min_in) {
                min_in = vs[i].indeg;
                min_in_ex = vs[i].exdeg;
            } else if (vs[i].indeg == min_in && vs[i].exdeg < min_in_ex) {
                min_in_ex = vs[i].exdeg;
            }
            int tt = vs[i].indeg + vs[i].exdeg;
            if (tt < min_tot) min_tot = tt;
        }
        med = get_mindeg((long)m + 1);
        if (min_in < md(m)) {
            st.lu_low++;
            lb = get_mindeg(m) - min_in;
            while (lb <= min_in_ex && min_in + lb < md((long)m + lb)) lb++;
            if (min_in + lb < md((long)m + lb)) {
                lb = nc + 1;
                invalid = true;
                st.lu_lb_invalid++;
            }
            ub = (int)std::floor((double)min_tot / gamma) + 1 - m;
            if (ub > nc) ub = nc;
            if (lb < ub) {
                long sc = 0;
                int idx;
                for (idx = 0; idx < lb; idx++) sc += ci[idx];
                while (idx < ub && sum_in + sc < (long)m * md((long)m + idx)) {
                    sc += ci[idx];
                    idx++;
                }
                if (sum_in + sc < (long)m * md((long)m + idx)) {
                    lb = ub + 1;
                    invalid = true;
                    st.lu_sum_invalid++;
                } else {
                    lb = idx;
                    int tub = idx;
                    while (idx < ub) {
                        sc += ci[idx];
                        idx++;
                        if (sum_in + sc >= (long)m * md((long)m + idx)) tub = idx;
                    }
                    if (ub > tub) {
                        ub = tub;
                        st.lu_ub_tightened++;
                    }
                    if (lb > 1) med = get_mindeg((long)m + lb);
                }
            }
        } else {
            ub = nc;
            lb = m < ms ? ms - m : 0;
        }
        if (m + ub < ms) invalid = true;
        if (ub < 0 || ub < lb) invalid = true;
        if (invalid) st.lu_invalid++;
        return invalid;
    }

    // ---- degree pruning, main.cu:1735-1855 ----
    // cind[0..nind) holds the indegrees of the surviving candidates.
    bool degree_pruning(Vtx* vs, int& t, int& c, int m, int& ub, int& lb, int& med, int nind) {
        std::sort(cind.begin(), cind.begin() + nind, [](int a, int b) { return a > b; });
        if (lu_bounds(ub, lb, med, vs, m, nind, cind.data())) return true;
        for (int k = 0; k < m; k++)
            if (!vert_ext(vs[k], m, ub, lb, med)) {
                st.member_fail++;
                return true;
            }
        int nrem = 0, nrmv = 0;
        for (int i = m; i < t; i++) {
            if (vs[i].label == 0 && cand_valid(vs[i], m, ub, lb, med))
                remaining[nrem++] = i;
            else
                removed[nrmv++] = i;
        }
        while (nrem > 0 && nrmv > 0) {
            if (nrem < nrmv) {
                st.dp_recompute++;
                for (int i = 0; i < t; i++) vs[i].exdeg = 0;
                for (int i = 0; i < nrem; i++) {
                    int v = vs[remaining[i]].id;
                    for (uint64_t j = roff[v]; j < roff[v + 1]; j++) {
                        int p = look(radj[j], t);
                        if (p > -1) vs[p].exdeg++;
                    }
                    st.w_adj += roff[v + 1] - roff[v];
                }
            } else {
                st.dp_decrement++;
                for (int i = 0; i < nrmv; i++) {
                    int v = vs[removed[i]].id;
                    for (uint64_t j = roff[v]; j < roff[v + 1]; j++) {
                        int p = look(radj[j], t);
                        if (p > -1) vs[p].exdeg--;
                    }
                    st.w_adj += roff[v + 1] - roff[v];
                }
            }
            int nv = 0;
            for (int k = 0; k < nrem; k++)
                if (cand_valid(vs[remaining[k]], m, ub, lb, med)) cind[nv++] = vs[remaining[k]].indeg;
            std::sort(cind.begin(), cind.begin() + nv, [](int a, int b) { return a > b; });
            if (lu_bounds(ub, lb, med, vs, m, nv, cind.data())) return true;
            for (int k = 0; k < m; k++)
                if (!vert_ext(vs[k], m, ub, lb, med)) {
                    st.member_fail++;
                    return true;
                }
            nv = 0;
            nrmv = 0;
            for (int k = 0; k < nrem; k++) {
                if (cand_valid(vs[remaining[k]], m, ub, lb, med))
                    remaining[nv++] = remaining[k];
                else
                    removed[nrmv++] = remaining[k];
            }
            nrem = nv;
        }
        for (int i = 0; i < nrem; i++) vs[m + i] = vs[remaining[i]];
        t = t - c + nrem;
        c = nrem;
        return false;
    }

    // ---- diameter pruning, main.cu:1709-1732 ----
    int diameter_pruning(Vtx* vs, int v, int t, int m) {
        int nind = 0;
        for (int i = m; i < t; i++) vs[i].label = -1;
        for (uint64_t j = r2off[v]; j < r2off[v + 1]; j++) {
            int p = look(r2adj[j], t);
            if (p >= m) {
                vs[p].label = 0;
                cind[nind++] = vs[p].indeg;
            }
        }
        st.w_adj += r2off[v + 1] - r2off[v];
        return nind;
    }

    // ---- add one vertex, main.cu:1465-1523 ----
    int add_one_vertex(Vtx* vs, int& t, int& c, int& m, int& ub, int& lb, int& med) {
        st.branches++;
        int v = vs[m].id;
        vs[m].label = 1;
        m++;
        c--;
        for (int i = 0; i < t; i++) mset(vs[i].id, i);
        for (uint64_t j = roff[v]; j < roff[v + 1]; j++) {
            int p = look(radj[j], t);
            if (p > -1) {
                vs[p].indeg++;
                vs[p].exdeg--;
            }
        }
        st.w_adj += roff[v + 1] - roff[v];
        int nind = diameter_pruning(vs, v, t, m);
        bool r = degree_pruning(vs, t, c, m, ub, lb, med, nind);
        full_reset();
        return r ? 1 : 0;
    }

    // ---- critical vertex pruning, main.cu:1526-1707 ----
    int critical_vertex_pruning(Vtx* vs, int& t, int& c, int& m, int& ub, int& lb, int& med) {
        for (int i = 0; i < t; i++) mset(vs[i].id, i);
        adjc.assign(t, 0);
        int thr = md((long)m + lb);
        for (int k = 0; k < m; k++) {
            if (vs[k].indeg + vs[k].exdeg == thr && vs[k].exdeg > 0) {
                int v = vs[k].id;
                for (uint64_t j = roff[v]; j < roff[v + 1]; j++) {
                    int p = look(radj[j], t);
                    if (p >= m) vs[p].label = 4;
                }
                st.w_adj += roff[v + 1] - roff[v];
            }
        }
        for (int i = 0; i < t; i++) munset(vs[i].id);
        // glibc qsort is a stable merge sort at these sizes; the order is a
        // partial order (label 4 first), so a stable sort reproduces it.
        std::stable_sort(vs + m, vs + t, [](const Vtx& a, const Vtx& b) {
            return a.label == 4 && b.label != 4;
        });
        int ncrit = 0;
        for (int i = m; i < t; i++) {
            if (vs[i].label == 4)
                ncrit++;
            else
                break;
        }
        if (ncrit > 0) {
            for (int i = 0; i < t; i++) mset(vs[i].id, i);
            for (int i = m; i < m + ncrit; i++) {
                int v = vs[i].id;
                for (uint64_t j = roff[v]; j < roff[v + 1]; j++) {
                    int p = look(radj[j], t);
                    if (p > -1) {
                        vs[p].indeg++;
                        vs[p].exdeg--;
                    }
                }
                for (uint64_t j = r2off[v]; j < r2off[v + 1]; j++) {
                    int p = look(r2adj[j], t);
                    if (p > -1) adjc[p]++;
                }
                st.w_adj += (roff[v + 1] - roff[v]) + (r2off[v + 1] - r2off[v]);
            }
            bool fail = false;
            for (int k = 0; k < m; k++)
                if (adjc[k] != ncrit) fail = true;
            if (!fail)
                for (int k = m; k < m + ncrit; k++)
                    if (adjc[k] < ncrit - 1) fail = true;
            if (fail) {
                st.crit_fail++;
                for (int i = 0; i < t; i++) munset(vs[i].id);
                if (!faithful) full_reset();
                return 2;
            }
            st.crit_ok++;
            for (int k = m; k < m + ncrit; k++) vs[k].label = 1;
            m += ncrit;
            c -= ncrit;
        } else {
            st.crit_none++;
        }
        int nind = 0;
        for (int k = m; k < t; k++) {
            if (adjc[k] == ncrit)
                cind[nind++] = vs[k].indeg;
            else
                vs[k].label = -1;
        }
        bool r = degree_pruning(vs, t, c, m, ub, lb, med, nind);
        for (int i = 0; i < t; i++) munset(vs[i].id);
        if (!faithful) full_reset();
        return r ? 1 : 0;
    }

    // ---- clique test, main.cu:1974-1995 ----
    void check_for_clique(const Vtx* vs, int m) {
        int req = md(m);
        for (int k = 0; k < m; k++)
            if (vs[k].indeg < req) return;
        for (int k = 0; k < m; k++) clq_v.push_back(vs[k].id);
        clq_off.push_back(clq_v.size());
        st.raw_cliques++;
        st.sum_clique_len += m;
    }

    // ---- lookahead, main.cu:1349-1405 ----
    int lookahead(Vtx* rv, int t, int nm, int nc) {
        for (int i = 0; i < nm; i++)
            if (rv[i].indeg + rv[i].exdeg < md(t)) return 0;
        for (int i = nm; i < t; i++) mset(rv[i].id, i);
        for (int i = nm; i < t; i++) {
            int v = rv[i].id;
            for (uint64_t j = r2off[v]; j < r2off[v + 1]; j++) {
                int p = look(r2adj[j], t);
                if (p >= nm) rv[p].l2++;
            }
            st.w_adj += r2off[v + 1] - r2off[v];
        }
        for (int i = nm; i < t; i++) munset(rv[i].id);
        for (int j = nm; j < t; j++)
            if (rv[j].l2 < nc - 1 || rv[j].indeg + rv[j].exdeg < md(t)) return 0;
        for (int j = 0; j < t; j++) clq_v.push_back(rv[j].id);
        clq_off.push_back(clq_v.size());
        st.raw_cliques++;
        st.sum_clique_len += t;
        st.lookahead_hits++;
        return 1;
    }

    // ---- remove one vertex, main.cu:1408-1462 ----
    int remove_one_vertex(Vtx* rv, int& t, int& nc, int nm) {
        int mdg = get_mindeg(nm);
        nc--;
        t--;
        for (int i = 0; i < t; i++) mset(rv[i].id, i);
        bool failed = false;
        int v = rv[t].id;
        uint64_t j;
        for (j = roff[v]; j < roff[v + 1]; j++) {
            int p = look(radj[j], t);
            if (p > -1) {
                rv[p].exdeg--;
                if (p < nm && rv[p].indeg + rv[p].exdeg < mdg) {
                    failed = true;
                    j++;
                    break;
                }
            }
        }
        st.w_adj += j - roff[v];
        for (int i = 0; i < t; i++) munset(rv[i].id);
        return failed ? 1 : 0;
    }

    // ---- one task record, main.cu:1060-1181 ----
    template <class EmitChild>
    void expand_task(std::vector<Vtx>& rec, EmitChild&& emit_child) {
        int t = (int)rec.size();
        int nm = 0;
        while (nm < t && rec[nm].label == 1) nm++;
        int cov = 0;
        while (nm + cov < t && rec[nm + cov].label == 2) cov++;
        int nc = t - nm;
        int expansions = nc;
        st.tasks++;
        st.sum_t += t;
        if ((uint64_t)t > st.max_t) st.max_t = t;
        if (!faithful) full_reset();
        if (lookahead(rec.data(), t, nm, nc)) return;
        std::vector<Vtx> vs;
        const bool root = cur_depth == 0;
        if (root) st.root_branches = expansions - cov;
        // the window of this task's branches: the root's, a depth-1 task's,
        // or all; a negative start counts from the end of the branches
        const bool win = cur_depth <= 1;
        const long wb0 = cur_depth == 0 ? rb0 : lb0, wb1 = cur_depth == 0 ? rb1 : lb1;
        const long b0 = !win ? 0 : wb0 >= 0 ? wb0 : std::max(0L, (long)(expansions - cov) + wb0);
        for (int j = cov; j < expansions; j++) {
            if (win && wb1 >= 0 && j - cov >= wb1) break;
            if (j != cov) {
                if (remove_one_vertex(rec.data(), t, nc, nm)) {
                    st.remove_breaks++;
                    break;
                }
            }
            if (j - cov < b0) continue;
            vs.resize(t);
            int m = nm, c = nc, tt = t;
            for (int i = 0; i < m; i++) vs[i] = rec[i];
            vs[m] = rec[t - 1];
            for (int i = m; i < t - 1; i++) vs[i + 1] = rec[i];
            if (cov > 0)
                for (int k = nm + 1; k <= nm + cov; k++) vs[k].label = 0;
            int ub = 0, lb = 0, med = 0;
            int r = add_one_vertex(vs.data(), tt, c, m, ub, lb, med);
            if (r == 1) {
                if (m >= ms) check_for_clique(vs.data(), m);
                continue;
            }
            r = critical_vertex_pruning(vs.data(), tt, c, m, ub, lb, med);
            if (r == 2) continue;
            if (m >= ms) check_for_clique(vs.data(), m);
            if (r == 1) continue;
            std::sort(vs.begin(), vs.begin() + tt, q_less);
            if (c > 0) {
                std::vector<Vtx> child(vs.begin(), vs.begin() + tt);
                for (auto& x : child) x.l2 = 0;
                st.children++;
                st.sum_t_written += tt;
                emit_child(std::move(child), ub - lb);
            }
        }
    }

    // |N2(v)| over the full graph (adjToSer.cpp:115-156), with an optional
    // early exit once the count reaches `cap` (cap < 0: exact).
    long n2_size(int v, std::vector<int>& mark, int stamp, long cap) {
        if (off2) return (long)(off2[v + 1] - off2[v]);
        long cnt = 0;
        mark[v] = stamp;
        for (uint64_t j = off[v]; j < off[v + 1]; j++) {
            int u = adj[j];
            if (mark[u] != stamp) {
                mark[u] = stamp;
                cnt++;
            }
            for (uint64_t k = off[u]; k < off[u + 1]; k++) {
                int w = adj[k];
                if (mark[w] != stamp) {
                    mark[w] = stamp;
                    cnt++;
                }
            }
            if (cap >= 0 && cnt >= cap) return cnt;
        }
        return cnt;
    }

    // ---- root task, main.cu:846-998 ----
    std::vector<Vtx> build_root() {
        int thr = md(ms);
        std::vector<Vtx> vs(n);
        std::vector<int> rem, rmv;
        std::vector<int> mark(n, -1);
        int stamp = 0;
        for (int i = 0; i < n; i++) {
            int deg = (int)(off[i + 1] - off[i]);
            vs[i].id = i;
            vs[i].indeg = 0;
            vs[i].exdeg = deg;
            vs[i].l2 = 0;
            bool pass = deg >= thr;
            if (pass && deg < ms - 1) pass = n2_size(i, mark, stamp++, ms - 1) >= ms - 1;
            if (pass) {
                vs[i].label = 0;
                rem.push_back(i);
            } else {
                vs[i].label = -1;
                rmv.push_back(i);
            }
        }
        int nc = n;
        while ((int)rem.size() < nc / 2) {
            nc = (int)rem.size();
            for (int v : rem) vs[v].exdeg = 0;
            for (int v : rem) {
                for (uint64_t j = off[v]; j < off[v + 1]; j++)
                    if (vs[adj[j]].label == 0) vs[adj[j]].exdeg++;
                st.init_w_adj += off[v + 1] - off[v];
            }
            std::vector<int> keep;
            rmv.clear();
            for (int v : rem) {
                if (vs[v].exdeg >= thr)
                    keep.push_back(v);
                else {
                    vs[v].label = -1;
                    rmv.push_back(v);
                }
            }
            rem.swap(keep);
        }
        nc = (int)rem.size();
        for (size_t k = 0; k < rmv.size(); k++) {
            int v = rmv[k];
            for (uint64_t j = off[v]; j < off[v + 1]; j++) {
                int u = adj[j];
                if (vs[u].label == 0) {
                    vs[u].exdeg--;
                    if (vs[u].exdeg < thr) {
                        vs[u].label = -1;
                        nc--;
                        rmv.push_back(u);
                    }
                }
            }
            st.init_w_adj += off[v + 1] - off[v];
        }
        // exact |N2| of survivors: the third sort key
        for (int i = 0; i < n; i++)
            if (vs[i].label == 0) vs[i].l2 = (int)n2_size(i, mark, stamp++, -1);
        int maxd = 0, maxi = 0;
        for (int i = 0; i < n; i++)
            if (vs[i].label == 0 && vs[i].exdeg > maxd) {
                maxd = vs[i].exdeg;
                maxi = i;
            }
        vs[maxi].label = 3;
        for (uint64_t j = off[maxi]; j < off[maxi + 1]; j++)
            if (vs[adj[j]].label == 0) vs[adj[j]].label = 2;
        std::sort(vs.begin(), vs.end(), q_less);
        vs.resize(nc);
        for (auto& x : vs) x.l2 = 0;
        return vs;
    }

    // restricted N / N2 lists over the root vertex set
    void build_restricted(const std::vector<Vtx>& root) {
        in_root.assign(n, 0);
        for (auto& x : root) in_root[x.id] = 1;
        roff_s.assign(n + 1, 0);
        r2off_s.assign(n + 1, 0);
        radj_s.clear();
        r2adj_s.clear();
        std::vector<int> mark(n, -1), buf;
        for (int v = 0; v < n; v++) {
            if (in_root[v]) {
                for (uint64_t j = off[v]; j < off[v + 1]; j++)
                    if (in_root[adj[j]]) radj_s.push_back(adj[j]);
                buf.clear();
                if (off2) {
                    for (uint64_t j = off2[v]; j < off2[v + 1]; j++)
                        if (in_root[adj2[j]]) buf.push_back(adj2[j]);
                } else {
                    mark[v] = v;
                    for (uint64_t j = off[v]; j < off[v + 1]; j++) {
                        int u = adj[j];
                        if (mark[u] != v) {
                            mark[u] = v;
                            if (in_root[u]) buf.push_back(u);
                        }
                        for (uint64_t k = off[u]; k < off[u + 1]; k++) {
                            int w = adj[k];
                            if (mark[w] != v) {
                                mark[w] = v;
                                if (in_root[w]) buf.push_back(w);
                            }
                        }
                    }
                }
                std::sort(buf.begin(), buf.end());
                r2adj_s.insert(r2adj_s.end(), buf.begin(), buf.end());
            }
            roff_s[v + 1] = radj_s.size();
            r2off_s[v + 1] = r2adj_s.size();
        }
        roff = roff_s.data();
        r2off = r2off_s.data();
        radj = radj_s.data();
        r2adj = r2adj_s.data();
    }

    // Task features for the cost model (12 values), computed from a record
    // when it is written: members, candidates, min/max/mean indeg and exdeg
    // over all vertices, depth, ub - lb of the bounds that admitted the
    // record (candidates for the root), mean |N2 n root| over candidates,
    // and the record size.  Means are float32(sum) / float32(count).
    std::vector<int> d2r;
    void features(const std::vector<Vtx>& rec, int depth, int gap, float* f) {
        int t = (int)rec.size();
        int nm = 0;
        while (nm < t && rec[nm].label == 1) nm++;
        int mi = INT32_MAX, ma = INT32_MIN, me = INT32_MAX, mx = INT32_MIN;
        long long si = 0, se = 0, s2 = 0;
        for (int k = 0; k < t; k++) {
            const Vtx& x = rec[k];
            mi = std::min(mi, x.indeg);
            ma = std::max(ma, x.indeg);
            si += x.indeg;
            me = std::min(me, x.exdeg);
            mx = std::max(mx, x.exdeg);
            se += x.exdeg;
            if (k >= nm) s2 += d2r[x.id];
        }
        int c = t - nm;
        f[0] = (float)nm;
        f[1] = (float)c;
        f[2] = (float)mi;
        f[3] = (float)ma;
        f[4] = (float)si / (float)t;
        f[5] = (float)me;
        f[6] = (float)mx;
        f[7] = (float)se / (float)t;
        f[8] = (float)depth;
        f[9] = (float)gap;
        f[10] = c > 0 ? (float)s2 / (float)c : 0.0f;
        f[11] = (float)t;
    }

    int run() {
        pos.assign(n, -1);
        remaining.assign(n + 1, 0);
        removed.assign(n + 1, 0);
        cind.assign(n + 1, 0);
        std::vector<Vtx> root = build_root();
        st.root_size = root.size();
        if (root.empty()) return 0;
        build_restricted(root);
        struct Item {
            std::vector<Vtx> rec;
            int depth;
            int64_t tid;
            float f[12];
        };
        if (trace) {
            d2r.assign(n, 0);
            for (int v = 0; v < n; v++) d2r[v] = (int)(r2off[v + 1] - r2off[v]);
        }
        std::vector<Item> stack;
        {
            Item r0{std::move(root), 0, -1, {}};
            if (trace) {
                features(r0.rec, 0, (int)r0.rec.size(), r0.f);
            }
            stack.push_back(std::move(r0));
        }
        std::vector<Item> kids;
        while (!stack.empty()) {
            Item it = std::move(stack.back());
            stack.pop_back();
            if ((uint64_t)it.depth > st.max_depth) st.max_depth = it.depth;
            if (max_depth >= 0 && it.depth > max_depth) {
                st.beyond++;
                if (trace && it.tid >= 0) trace_rows[it.tid].beyond++;
                continue;
            }
            int64_t my = -1;
            if (trace) {
                TraceRow r;
                memcpy(r.f, it.f, sizeof(r.f));
                r.parent = it.tid;
                r.beyond = 0;
                my = (int64_t)trace_rows.size();
                trace_rows.push_back(r);
            }
            cur_depth = it.depth;
            kids.clear();
            int d = it.depth + 1;
            expand_task(it.rec, [&](std::vector<Vtx>&& c, int gap) {
                Item k{std::move(c), d, my, {}};
                if (trace) features(k.rec, d, gap, k.f);
                kids.push_back(std::move(k));
            });
            // children are processed in branch order (DFS preorder): the
            // first child written by the root is the next task processed,
            // as in the reference's second host level.
            for (auto k = kids.rbegin(); k != kids.rend(); ++k) stack.push_back(std::move(*k));
        }
        return 0;
    }

    // The same search on `threads` host threads (clean map mode, no trace):
    // tasks are independent once the map starts clean, so workers take task
    // records from a shared pool, run their subtrees depth first and hand
    // records back to the pool whenever it runs low.  The counters are sums
    // over tasks and the raw sets a multiset, so both equal the sequential
    // run's; only the order of the raw sets differs.
    int run_mt(int threads) {
        std::vector<Vtx> root = build_root();
        st.root_size = root.size();
        if (root.empty()) return 0;
        build_restricted(root);
        struct Item {
            std::vector<Vtx> rec;
            int depth;
        };
        std::mutex mu;
        std::condition_variable cv;
        std::deque<Item> pool;
        int busy = 0;
        pool.push_back(Item{std::move(root), 0});
        std::vector<Oracle> ws(threads);
        for (Oracle& w : ws) {
            w.n = n;
            w.off = off;
            w.adj = adj;
            w.off2 = off2;
            w.adj2 = adj2;
            w.roff = roff;
            w.r2off = r2off;
            w.radj = radj;
            w.r2adj = r2adj;
            w.gamma = gamma;
            w.ms = ms;
            w.mindeg = mindeg;
            w.rb0 = rb0;
            w.rb1 = rb1;
            w.lb0 = lb0;
            w.lb1 = lb1;
            w.max_depth = max_depth;
            w.pos.assign(n, -1);
            w.remaining.assign(n + 1, 0);
            w.removed.assign(n + 1, 0);
            w.cind.assign(n + 1, 0);
        }
        auto work = [&](Oracle& w) {
            std::vector<Item> stack, kids;
            for (;;) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return !pool.empty() || busy == 0; });
                    if (pool.empty()) return;
                    stack.push_back(std::move(pool.front()));
                    pool.pop_front();
                    busy++;
                }
                while (!stack.empty()) {
                    Item it = std::move(stack.back());
                    stack.pop_back();
                    if ((uint64_t)it.depth > w.st.max_depth) w.st.max_depth = it.depth;
                    if (max_depth >= 0 && it.depth > max_depth) {
                        w.st.beyond++;
                        continue;
                    }
                    w.cur_depth = it.depth;
                    kids.clear();
                    const int d = it.depth + 1;
                    w.expand_task(it.rec, [&](std::vector<Vtx>&& c, int) { kids.push_back(Item{std::move(c), d}); });
                    for (auto k = kids.rbegin(); k != kids.rend(); ++k) stack.push_back(std::move(*k));
                    // share the oldest (shallowest) records while others wait
                    if (stack.size() > 1) {
                        std::lock_guard<std::mutex> lk(mu);
                        size_t give = 0;
                        while (pool.size() < (size_t)threads && stack.size() > 1 && give < stack.size() / 2) {
                            pool.push_back(std::move(stack.front()));
                            stack.erase(stack.begin());
                            give++;
                        }
                        if (give) cv.notify_all();
                    }
                }
                {
                    std::lock_guard<std::mutex> lk(mu);
                    busy--;
                    if (busy == 0 && pool.empty()) cv.notify_all();
                }
            }
        };
        std::vector<std::thread> th;
        for (int i = 1; i < threads; i++) th.emplace_back(work, std::ref(ws[i]));
        work(ws[0]);
        for (auto& t : th) t.join();
        for (Oracle& w : ws) {
            st.tasks += w.st.tasks;
            st.branches += w.st.branches;
            st.children += w.st.children;
            st.raw_cliques += w.st.raw_cliques;
            st.sum_t += w.st.sum_t;
            st.sum_t_written += w.st.sum_t_written;
            st.sum_clique_len += w.st.sum_clique_len;
            st.w_adj += w.st.w_adj;
            st.lookahead_hits += w.st.lookahead_hits;
            st.stale_oob += w.st.stale_oob;
            st.stale_hits += w.st.stale_hits;
            st.max_t = std::max(st.max_t, w.st.max_t);
            st.max_depth = std::max(st.max_depth, w.st.max_depth);
            st.beyond += w.st.beyond;
            st.root_branches += w.st.root_branches;
            st.remove_breaks += w.st.remove_breaks;
            st.lu_low += w.st.lu_low;
            st.lu_lb_invalid += w.st.lu_lb_invalid;
            st.lu_sum_invalid += w.st.lu_sum_invalid;
            st.lu_ub_tightened += w.st.lu_ub_tightened;
            st.lu_invalid += w.st.lu_invalid;
            st.member_fail += w.st.member_fail;
            st.dp_recompute += w.st.dp_recompute;
            st.dp_decrement += w.st.dp_decrement;
            st.crit_none += w.st.crit_none;
            st.crit_fail += w.st.crit_fail;
            st.crit_ok += w.st.crit_ok;
            for (size_t i = 0; i + 1 < w.clq_off.size(); i++) {
                clq_v.insert(clq_v.end(), w.clq_v.begin() + w.clq_off[i], w.clq_v.begin() + w.clq_off[i + 1]);
                clq_off.push_back(clq_v.size());
            }
        }
        return 0;
    }
};

}  // namespace

// ----------------------------------------------------------------------------
// C interface (ctypes)
// ----------------------------------------------------------------------------
extern "C" {

struct orc_stats {
    uint64_t tasks, branches, children, raw_cliques, sum_t, sum_t_written, sum_clique_len, w_adj,
        lookahead_hits, stale_oob, stale_hits, max_t, max_depth, root_size, init_w_adj, beyond, root_branches,
        remove_breaks, lu_low, lu_lb_invalid, lu_sum_invalid, lu_ub_tightened, lu_invalid, member_fail,
        dp_recompute, dp_decrement, crit_none, crit_fail, crit_ok;
};

struct orc_result {
    std::vector<uint64_t> off;
    std::vector<int> v;
    orc_stats st;
    std::vector<float> trace_f;
    std::vector<int64_t> trace_parent;
    std::vector<int64_t> trace_beyond;
};

// Mine a graph to completion.  off2/adj2 may be NULL (N2 derived from N).
orc_result* orc_mine_window(int n, const uint64_t* off, const int* adj, const uint64_t* off2, const int* adj2,
                            double gamma, int min_size, int faithful, int trace, long rb0, long rb1, int max_depth);
orc_result* orc_mine_window_mt(int n, const uint64_t* off, const int* adj, const uint64_t* off2, const int* adj2,
                               double gamma, int min_size, int faithful, int trace, long rb0, long rb1, int max_depth,
                               int threads);
orc_result* orc_mine_window2(int n, const uint64_t* off, const int* adj, const uint64_t* off2, const int* adj2,
                             double gamma, int min_size, int faithful, int trace, long rb0, long rb1, long lb0,
                             long lb1, int max_depth, int threads);

orc_result* orc_mine(int n, const uint64_t* off, const int* adj, const uint64_t* off2, const int* adj2,
                     double gamma, int min_size, int faithful, int trace) {
    return orc_mine_window(n, off, adj, off2, adj2, gamma, min_size, faithful, trace, 0, -1, -1);
}

// Same, restricted to root branches [rb0, rb1) (rb1 < 0: all; rb0 < 0: the
// last -rb0 branches) and to tasks of depth <= max_depth (max_depth < 0: all).
orc_result* orc_mine_window(int n, const uint64_t* off, const int* adj, const uint64_t* off2, const int* adj2,
                            double gamma, int min_size, int faithful, int trace, long rb0, long rb1, int max_depth) {
    return orc_mine_window_mt(n, off, adj, off2, adj2, gamma, min_size, faithful, trace, rb0, rb1, max_depth, 1);
}

// threads > 1: the multi-threaded search (clean map mode and no trace only)
orc_result* orc_mine_window_mt(int n, const uint64_t* off, const int* adj, const uint64_t* off2, const int* adj2,
                               double gamma, int min_size, int faithful, int trace, long rb0, long rb1, int max_depth,
                               int threads) {
    return orc_mine_window2(n, off, adj, off2, adj2, gamma, min_size, faithful, trace, rb0, rb1, 0, -1, max_depth,
                            threads);
}

// ... and the branches [lb0, lb1) of every depth-1 task
orc_result* orc_mine_window2(int n, const uint64_t* off, const int* adj, const uint64_t* off2, const int* adj2,
                             double gamma, int min_size, int faithful, int trace, long rb0, long rb1, long lb0,
                             long lb1, int max_depth, int threads) {
    if (threads > 1 && (faithful || trace)) return nullptr;
    Oracle o;
    o.n = n;
    o.off = off;
    o.adj = adj;
    o.off2 = off2;
    o.adj2 = adj2;
    o.gamma = gamma;
    o.ms = min_size;
    o.faithful = faithful != 0;
    o.trace = trace != 0;
    o.rb0 = rb0;
    o.rb1 = rb1;
    o.lb0 = lb0;
    o.lb1 = lb1;
    o.max_depth = max_depth;
    o.build_tables();
    if (threads > 1)
        o.run_mt(threads);
    else
        o.run();
    orc_result* r = new orc_result;
    r->off = std::move(o.clq_off);
    r->v = std::move(o.clq_v);
    static_assert(sizeof(orc_stats) == sizeof(Stats), "stats layout");
    memcpy(&r->st, &o.st, sizeof(Stats));
    if (o.trace) {
        r->trace_f.resize(o.trace_rows.size() * 12);
        r->trace_parent.resize(o.trace_rows.size());
        r->trace_beyond.resize(o.trace_rows.size());
        for (size_t i = 0; i < o.trace_rows.size(); i++) {
            memcpy(&r->trace_f[i * 12], o.trace_rows[i].f, 12 * sizeof(float));
            r->trace_parent[i] = o.trace_rows[i].parent;
            r->trace_beyond[i] = o.trace_rows[i].beyond;
        }
    }
    return r;
}

// The root task record of initialize_tasks (main.cu:846-998) in its Q
// order: ids, labels (0 candidate, 2 covered, 3 cover) and exdegs.  Returns
// the record size; the arrays (n entries each) may be NULL.
int orc_root(int n, const uint64_t* off, const int* adj, double gamma, int min_size, int* ids, int* labels,
             int* exdeg) {
    Oracle o;
    o.n = n;
    o.off = off;
    o.adj = adj;
    o.gamma = gamma;
    o.ms = min_size;
    o.build_tables();
    std::vector<Vtx> root = o.build_root();
    for (size_t i = 0; i < root.size(); i++) {
        if (ids) ids[i] = root[i].id;
        if (labels) labels[i] = root[i].label;
        if (exdeg) exdeg[i] = root[i].exdeg;
    }
    return (int)root.size();
}

// lu_bounds (main.cu:1857-1972) on a given record: members' (indeg,
// exdeg) and the candidates' indegrees sorted descending.  out = {invalid,
// lb, ub, min_ext_deg}.
void orc_lu_bounds(double gamma, int min_size, int m, const int* indeg, const int* exdeg, int nc, const int* cind,
                   int* out) {
    Oracle o;
    o.n = 2 * (m + nc) + 16;
    o.gamma = gamma;
    o.ms = min_size;
    o.build_tables();
    std::vector<Vtx> vs(m);
    for (int i = 0; i < m; i++) vs[i] = Vtx{i, 1, indeg[i], exdeg[i], 0};
    int ub = 0, lb = 0, med = 0;
    out[0] = o.lu_bounds(ub, lb, med, vs.data(), m, nc, cind) ? 1 : 0;
    out[1] = lb;
    out[2] = ub;
    out[3] = med;
}

uint64_t orc_count(const orc_result* r) { return r->off.size() - 1; }
const uint64_t* orc_off(const orc_result* r) { return r->off.data(); }
const int* orc_verts(const orc_result* r) { return r->v.data(); }
void orc_get_stats(const orc_result* r, orc_stats* s) { *s = r->st; }
uint64_t orc_trace_rows(const orc_result* r) { return r->trace_parent.size(); }
const float* orc_trace_feats(const orc_result* r) { return r->trace_f.data(); }
const int64_t* orc_trace_parent(const orc_result* r) { return r->trace_parent.data(); }
const int64_t* orc_trace_beyond(const orc_result* r) { return r->trace_beyond.data(); }
void orc_free(orc_result* r) { delete r; }

// RemoveNonMax semantics of main.cu:4569-4961: each set is sorted ascending,
// duplicates merge, every set that is a proper subset of another is dropped,
// and the rest come out in lexicographic order (the trie DFS order of
// OutputMaxSet).  Restated with an inverted index instead of a trie.
orc_result* orc_nonmax(uint64_t k, const uint64_t* off, const int* v) {
    std::vector<std::vector<int>> sets(k);
    for (uint64_t i = 0; i < k; i++) {
        sets[i].assign(v + off[i], v + off[i + 1]);
        std::sort(sets[i].begin(), sets[i].end());
    }
    std::sort(sets.begin(), sets.end());
    sets.erase(std::unique(sets.begin(), sets.end()), sets.end());
    int maxv = -1;
    for (auto& s : sets)
        for (int x : s) maxv = std::max(maxv, x);
    std::vector<std::vector<uint32_t>> post(maxv + 1);
    for (size_t i = 0; i < sets.size(); i++)
        for (int x : sets[i]) post[x].push_back((uint32_t)i);
    orc_result* r = new orc_result;
    r->off.push_back(0);
    memset(&r->st, 0, sizeof(r->st));
    for (size_t i = 0; i < sets.size(); i++) {
        const auto& s = sets[i];
        int rare = s[0];
        for (int x : s)
            if (post[x].size() < post[rare].size()) rare = x;
        bool sub = false;
        for (uint32_t j : post[rare]) {
            if (j == i || sets[j].size() <= s.size()) continue;
            if (std::includes(sets[j].begin(), sets[j].end(), s.begin(), s.end())) {
                sub = true;
                break;
            }
        }
        if (!sub) {
            r->v.insert(r->v.end(), s.begin(), s.end());
            r->off.push_back(r->v.size());
        }
    }
    return r;
}

// Full-graph N2 lists with adjToSer.cpp:115-156 semantics (sorted, self
// excluded).  Returns total entries; fills off2 (n+1) and, if adj2 != NULL,
// the lists.  Call once with adj2 == NULL to size the buffer.
uint64_t orc_build_n2(int n, const uint64_t* off, const int* adj, uint64_t* off2, int* adj2) {
    std::vector<int> mark(n, -1), buf;
    uint64_t tot = 0;
    off2[0] = 0;
    for (int v = 0; v < n; v++) {
        buf.clear();
        for (uint64_t j = off[v]; j < off[v + 1]; j++) {
            int u = adj[j];
            if (mark[u] != v) {
                mark[u] = v;
                buf.push_back(u);
            }
            for (uint64_t k = off[u]; k < off[u + 1]; k++) {
                int w = adj[k];
                if (w != v && mark[w] != v) {
                    mark[w] = v;
                    buf.push_back(w);
                }
            }
        }
        std::sort(buf.begin(), buf.end());
        if (adj2) memcpy(adj2 + tot, buf.data(), buf.size() * sizeof(int));
        tot += buf.size();
        off2[v + 1] = tot;
    }
    return tot;
}

// Brute force for tiny graphs (n <= 24): every maximal gamma-quasi-clique of
// size >= min_size, by definition (each member has >= ceil(gamma(|S|-1))
// neighbours inside S), sorted lexicographically.
orc_result* orc_brute(int n, const uint64_t* off, const int* adj, double gamma, int min_size) {
    orc_result* r = new orc_result;
    r->off.push_back(0);
    memset(&r->st, 0, sizeof(r->st));
    if (n > 24) return r;
    std::vector<uint32_t> nb(n, 0);
    for (int v = 0; v < n; v++)
        for (uint64_t j = off[v]; j < off[v + 1]; j++) nb[v] |= 1u << adj[j];
    uint32_t full = n == 32 ? 0xffffffffu : ((1u << n) - 1);
    std::vector<uint32_t> qcs;
    for (uint32_t s = 1; s <= full && s != 0; s++) {
        int sz = __builtin_popcount(s);
        if (sz < min_size) continue;
        int req = (int)std::ceil(gamma * (sz - 1));
        bool ok = true;
        for (uint32_t x = s; x && ok; x &= x - 1) {
            int v = __builtin_ctz(x);
            if (__builtin_popcount(nb[v] & s) < req) ok = false;
        }
        if (ok) qcs.push_back(s);
        if (s == full) break;
    }
    std::vector<std::vector<int>> out;
    for (uint32_t s : qcs) {
        bool maximal = true;
        for (uint32_t u : qcs)
            if (u != s && (u & s) == s) {
                maximal = false;
                break;
            }
        if (maximal) {
            std::vector<int> a;
            for (uint32_t x = s; x; x &= x - 1) a.push_back(__builtin_ctz(x));
            out.push_back(a);
        }
    }
    std::sort(out.begin(), out.end());
    for (auto& a : out) {
        r->v.insert(r->v.end(), a.begin(), a.end());
        r->off.push_back(r->v.size());
    }
    return r;
}

}  // extern "C"
