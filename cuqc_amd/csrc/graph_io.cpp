// graph_io.cpp -- graph construction, file formats and host tables.

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "qc_host.h"

namespace qc {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

int csr_from_edges(int n, std::vector<std::pair<uint32_t, uint32_t>>& edges, qc_graph* g) {
    g->n = n;
    g->off.assign((size_t)n + 1, 0);
    for (auto& e : edges) {
        if (e.first >= (uint32_t)n || e.second >= (uint32_t)n) {
            set_error("edge (%u,%u) out of range for n=%d", e.first, e.second, n);
            return QC_EINVAL;
        }
        if (e.first == e.second) continue;
        g->off[e.first + 1]++;
        g->off[e.second + 1]++;
    }
    for (int i = 0; i < n; i++) g->off[i + 1] += g->off[i];
    g->adj.assign(g->off[n], 0);
    std::vector<uint64_t> at(g->off.begin(), g->off.end() - 1);
    for (auto& e : edges) {
        if (e.first == e.second) continue;
        g->adj[at[e.first]++] = e.second;
        g->adj[at[e.second]++] = e.first;
    }
    // sort + dedupe each list, then compact
    uint64_t w = 0;
    uint64_t prev = 0;
    for (int i = 0; i < n; i++) {
        uint64_t b = prev, en = g->off[i + 1];
        std::sort(g->adj.begin() + b, g->adj.begin() + en);
        uint64_t start = w;
        for (uint64_t j = b; j < en; j++)
            if (j == b || g->adj[j] != g->adj[j - 1]) g->adj[w++] = g->adj[j];
        prev = en;
        g->off[i] = start;
    }
    g->off[n] = w;
    g->adj.resize(w);
    g->adj.shrink_to_fit();
    g->has_n2 = false;
    return QC_OK;
}

uint64_t fnv1a_graph(const qc_graph* g) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](const void* p, size_t len) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < len; i++) {
            h ^= b[i];
            h *= 1099511628211ULL;
        }
    };
    int32_t n = g->n;
    mix(&n, sizeof(n));
    mix(g->off.data(), g->off.size() * sizeof(uint64_t));
    mix(g->adj.data(), g->adj.size() * sizeof(uint32_t));
    return h;
}

void build_tables(double gamma, int n, int min_size, std::vector<int32_t>& mindeg, std::vector<int32_t>& ubq) {
    // Index ranges reached by the bound logic stay below 2 max(n, min_size)
    // + 2; the table is extended by the same formula beyond n.
    size_t sz = 2 * (size_t)std::max(n, min_size) + 16;
    mindeg.assign(sz, 0);
    for (size_t i = 1; i < sz; i++) mindeg[i] = (int32_t)std::ceil(gamma * (double)((int)i - 1));
    // min_clq_totaldeg <= n - 1
    ubq.assign((size_t)n + 2, 0);
    for (size_t d = 0; d < ubq.size(); d++) ubq[d] = (int32_t)std::floor((double)(int)d / gamma);
}

// ---------------------------------------------------------------------------
// text parsing helpers
// ---------------------------------------------------------------------------
namespace {

struct Text {
    std::vector<char> buf;
    size_t at = 0;
    bool load(const char* path) {
        FILE* f = fopen(path, "rb");
        if (!f) return false;
        fseek(f, 0, SEEK_END);
        long sz = ftell(f);
        fseek(f, 0, SEEK_SET);
        buf.resize(sz > 0 ? (size_t)sz + 1 : 1);
        size_t rd = sz > 0 ? fread(buf.data(), 1, sz, f) : 0;
        fclose(f);
        buf.resize(rd + 1);
        buf[rd] = 0;
        return true;
    }
    // next unsigned integer token; false at end or on a malformed token
    bool next(uint64_t& v) {
        const char* p = buf.data() + at;
        while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') p++;
        if (*p < '0' || *p > '9') {
            at = p - buf.data();
            return false;
        }
        uint64_t x = 0;
        while (*p >= '0' && *p <= '9') x = x * 10 + (uint64_t)(*p++ - '0');
        at = p - buf.data();
        v = x;
        return true;
    }
};

}  // namespace

int load_serialized(const char* path, qc_graph* g) {
    Text t;
    if (!t.load(path)) {
        set_error("invalid graph file: cannot open %s", path);
        return QC_EIO;
    }
    uint64_t V, E, L2;
    if (!t.next(V) || !t.next(E) || !t.next(L2) || V > (uint64_t)INT32_MAX) {
        set_error("%s: bad serialized header", path);
        return QC_EIO;
    }
    g->n = (int32_t)V;
    g->adj.resize(E);
    g->off.resize(V + 1);
    g->adj2.resize(L2);
    g->off2.resize(V + 1);
    uint64_t x;
    for (uint64_t i = 0; i < E; i++) {
        if (!t.next(x) || x >= V) goto bad;
        g->adj[i] = (uint32_t)x;
    }
    for (uint64_t i = 0; i <= V; i++) {
        if (!t.next(x) || x > E || (i > 0 && x < g->off[i - 1])) goto bad;
        g->off[i] = x;
    }
    if (g->off[0] != 0 || g->off[V] != E) goto bad;
    for (uint64_t i = 0; i < L2; i++) {
        if (!t.next(x) || x >= V) goto bad;
        g->adj2[i] = (uint32_t)x;
    }
    for (uint64_t i = 0; i <= V; i++) {
        if (!t.next(x) || x > L2 || (i > 0 && x < g->off2[i - 1])) goto bad;
        g->off2[i] = x;
    }
    if (g->off2[0] != 0 || g->off2[V] != L2) goto bad;
    g->has_n2 = true;
    return QC_OK;
bad:
    set_error("%s: short or corrupt serialized graph", path);
    return QC_EIO;
}

int load_adjlist(const char* path, qc_graph* g) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        set_error("invalid graph file: cannot open %s", path);
        return QC_EIO;
    }
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    std::vector<std::vector<uint32_t>> lists;
    std::vector<uint32_t> cur;
    int c;
    uint64_t num = 0;
    bool in_num = false;
    uint64_t maxv = 0;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
        if (c >= '0' && c <= '9') {
            num = num * 10 + (c - '0');
            in_num = true;
        } else {
            if (in_num) {
                cur.push_back((uint32_t)num);
                maxv = std::max(maxv, num);
                any = true;
            }
            num = 0;
            in_num = false;
            if (c == '\n') {
                lists.push_back(cur);
                cur.clear();
            } else if (c != ' ' && c != '\r' && c != '\t') {
                fclose(f);
                set_error("%s: unexpected character in adjacency list", path);
                return QC_EIO;
            }
        }
    }
    fclose(f);
    if (in_num) {
        cur.push_back((uint32_t)num);
        maxv = std::max(maxv, num);
        any = true;
    }
    if (!cur.empty()) lists.push_back(cur);
    size_t n = lists.size();
    if (any && maxv >= n) n = maxv + 1;
    for (size_t v = 0; v < lists.size(); v++)
        for (uint32_t u : lists[v]) edges.push_back({(uint32_t)v, u});
    int rc = csr_from_edges((int)n, edges, g);
    return rc;
}

int write_serialized(const qc_graph* g, const char* path) {
    FILE* f = fopen(path, "w");
    if (!f) {
        set_error("cannot open %s for write", path);
        return QC_EIO;
    }
    int n = g->n;
    // N2 with adjToSer.cpp:115-156 semantics
    std::vector<uint64_t> off2((size_t)n + 1, 0);
    std::vector<uint32_t> adj2;
    std::vector<int> mark(n, -1);
    std::vector<uint32_t> buf;
    for (int v = 0; v < n; v++) {
        buf.clear();
        for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
            uint32_t u = g->adj[j];
            if (mark[u] != v) {
                mark[u] = v;
                buf.push_back(u);
            }
            for (uint64_t k = g->off[u]; k < g->off[u + 1]; k++) {
                uint32_t w = g->adj[k];
                if ((int)w != v && mark[w] != v) {
                    mark[w] = v;
                    buf.push_back(w);
                }
            }
        }
        std::sort(buf.begin(), buf.end());
        adj2.insert(adj2.end(), buf.begin(), buf.end());
        off2[v + 1] = adj2.size();
    }
    uint64_t E = g->adj.size();
    fprintf(f, "%d\n%llu\n%llu\n", n, (unsigned long long)E, (unsigned long long)adj2.size());
    for (uint64_t i = 0; i < E; i++) fprintf(f, i + 1 < E ? "%u " : "%u", g->adj[i]);
    fputc('\n', f);
    for (int i = 0; i <= n; i++) fprintf(f, i < n ? "%llu " : "%llu", (unsigned long long)g->off[i]);
    fputc('\n', f);
    for (uint64_t i = 0; i < adj2.size(); i++) fprintf(f, i + 1 < adj2.size() ? "%u " : "%u", adj2[i]);
    fputc('\n', f);
    for (int i = 0; i <= n; i++) fprintf(f, i < n ? "%llu " : "%llu", (unsigned long long)off2[i]);
    fputc('\n', f);
    if (fclose(f) != 0) {
        set_error("write to %s failed", path);
        return QC_EIO;
    }
    return QC_OK;
}

int write_binary(const qc_graph* g, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_error("cannot open %s for write", path);
        return QC_EIO;
    }
    char magic[4] = {'Q', 'C', 'G', '1'};
    int32_t n = g->n;
    uint64_t nnz = g->adj.size();
    bool ok = fwrite(magic, 1, 4, f) == 4 && fwrite(&n, 4, 1, f) == 1 && fwrite(&nnz, 8, 1, f) == 1 &&
              fwrite(g->off.data(), 8, (size_t)n + 1, f) == (size_t)n + 1 &&
              fwrite(g->adj.data(), 4, nnz, f) == nnz;
    if (fclose(f) != 0 || !ok) {
        set_error("write to %s failed", path);
        return QC_EIO;
    }
    return QC_OK;
}

int load_binary(const char* path, qc_graph* g) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        set_error("invalid graph file: cannot open %s", path);
        return QC_EIO;
    }
    char magic[4];
    int32_t n;
    uint64_t nnz;
    bool ok = fread(magic, 1, 4, f) == 4 && memcmp(magic, "QCG1", 4) == 0 && fread(&n, 4, 1, f) == 1 &&
              fread(&nnz, 8, 1, f) == 1 && n >= 0;
    if (ok) {
        g->n = n;
        g->off.resize((size_t)n + 1);
        g->adj.resize(nnz);
        ok = fread(g->off.data(), 8, (size_t)n + 1, f) == (size_t)n + 1 && fread(g->adj.data(), 4, nnz, f) == nnz;
    }
    fclose(f);
    if (ok) {
        ok = g->off[0] == 0 && g->off[n] == nnz;
        for (int i = 0; ok && i < n; i++) ok = g->off[i] <= g->off[i + 1];
        for (uint64_t j = 0; ok && j < nnz; j++) ok = g->adj[j] < (uint32_t)n;
    }
    if (!ok) {
        set_error("%s: short or corrupt binary graph", path);
        return QC_EIO;
    }
    g->has_n2 = false;
    return QC_OK;
}

}  // namespace qc
