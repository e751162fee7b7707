// gen.cpp -- seeded synthetic graphs for the five benchmark workloads.
//
// All generators are deterministic functions of (config, seed, instance):
// RNG = SplitMix64-seeded xoshiro256**, uniform doubles from the top 53 bits,
// normals by Box-Muller.  Output: symmetric uint32 CSR, sorted lists, no
// self loops.
//
//   0  Erdos-Renyi G(200, 0.08), seed 1, planted QCs of sizes 20, 22, 25
//      (gamma 0.9), mined with gamma 0.9, min_size 10.
//   1  8,000 vertices with rank-16 Gaussian features; edge iff
//      |cos(f_i, f_j)| >= t, t the quantile giving 1,000,000 edges.  Features
//      follow a gene-module model: 32 modules with Gaussian centres mu_k and
//      Zipf sizes (share ~ 1/(k+1)) hold 90% of the vertices, a member is
//      mu_k + 0.3 z, the other 10% are isotropic; this gives the heavy-tailed
//      degrees of a co-expression network.  gamma 0.9, min_size 30.
//   2  Chung-Lu, n = 100,000, mean degree 20, exponent 2.1 (weights
//      w_i ~ (i + i0)^(-1/(beta-1)), capped at sqrt(sum w)), plus 50 planted
//      0.85-QCs of size U[20, 60]; gamma 0.85, min_size 20.
//   3  instance i of 256: ER or Chung-Lu (exponent U[2.1, 2.8]), n
//      log-uniform in [500, 20000], mean degree U[8, 60], 1-10 planted QCs of
//      size U[10, 40]; gamma U[0.6, 0.95] (2 decimals), min_size U[8, 40].
//   4  Chung-Lu, n = 1,000,000, mean degree 30, exponent 2.2, plus 200
//      planted 0.9-QCs of size U[30, 80]; gamma 0.9, min_size 40.
//
// Chung-Lu edges are drawn as M = n*d/2 independent endpoint pairs, each
// endpoint chosen with probability proportional to its weight (alias table);
// self loops and repeated pairs are dropped.
//
// Planting (every config): disjoint random member sets; every internal pair
// becomes an edge with probability 0.95; then, while some member has fewer
// than ceil(g (k-1)) internal neighbours, the member with the lowest internal
// degree (ties: lowest id) is joined to the non-adjacent member with the
// lowest internal degree (ties: lowest id).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "qc_host.h"

namespace qc {
namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) {
            x += 0x9e3779b97f4a7c15ULL;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (double)(next() >> 11) * 0x1.0p-53; }
    uint64_t below(uint64_t k) {  // unbiased in [0, k)
        uint64_t lim = UINT64_MAX - UINT64_MAX % k;
        uint64_t x;
        do x = next();
        while (x >= lim);
        return x % k;
    }
    bool have = false;
    double spare = 0;
    double normal() {
        if (have) {
            have = false;
            return spare;
        }
        double u1, u2;
        do u1 = uni();
        while (u1 <= 0.0);
        u2 = uni();
        double r = std::sqrt(-2.0 * std::log(u1));
        spare = r * std::sin(2.0 * M_PI * u2);
        have = true;
        return r * std::cos(2.0 * M_PI * u2);
    }
};

using Edges = std::vector<std::pair<uint32_t, uint32_t>>;

static inline uint64_t key(uint32_t a, uint32_t b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)a << 32) | b;
}

// G(n, p) by geometric skipping over the pairs i < j.
void er_edges(Rng& r, int n, double p, Edges& e) {
    if (p <= 0) return;
    double lq = std::log(1.0 - p);
    int64_t v = 1, w = -1;
    while (v < n) {
        double u = r.uni();
        w += 1 + (int64_t)std::floor(std::log(1.0 - u) / lq);
        while (w >= v && v < n) {
            w -= v;
            v++;
        }
        if (v < n) e.push_back({(uint32_t)w, (uint32_t)v});
    }
}

// Chung-Lu by weighted endpoint sampling with an alias table.
void chung_lu_edges(Rng& r, int n, double mean_deg, double beta, Edges& e) {
    std::vector<double> w(n);
    double a = 1.0 / (beta - 1.0);
    double i0 = 1.0;
    double sum = 0;
    for (int i = 0; i < n; i++) {
        w[i] = std::pow((double)i + i0, -a);
        sum += w[i];
    }
    double scale = mean_deg * n / sum;
    sum = 0;
    for (int i = 0; i < n; i++) {
        w[i] *= scale;
        sum += w[i];
    }
    double cap = std::sqrt(sum);
    sum = 0;
    for (int i = 0; i < n; i++) {
        if (w[i] > cap) w[i] = cap;
        sum += w[i];
    }
    // alias table (Vose)
    std::vector<double> prob(n);
    std::vector<uint32_t> alias(n);
    std::vector<uint32_t> small, large;
    for (int i = 0; i < n; i++) {
        prob[i] = w[i] * n / sum;
        (prob[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t s = small.back(), l = large.back();
        small.pop_back();
        alias[s] = l;
        prob[l] = prob[l] + prob[s] - 1.0;
        if (prob[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (uint32_t i : large) prob[i] = 1.0;
    for (uint32_t i : small) prob[i] = 1.0;
    auto draw = [&]() -> uint32_t {
        uint32_t i = (uint32_t)r.below(n);
        return r.uni() < prob[i] ? i : alias[i];
    };
    uint64_t m = (uint64_t)std::llround(mean_deg * n / 2.0);
    e.reserve(e.size() + m);
    for (uint64_t k = 0; k < m; k++) {
        uint32_t x = draw(), y = draw();
        if (x != y) e.push_back({x, y});
    }
}

// Plant `sizes.size()` quasi-cliques on disjoint random vertex sets; their
// member lists (ascending) are appended to `sets`.
void plant(Rng& r, int n, const std::vector<int>& sizes, double g, Edges& e, std::vector<std::vector<uint32_t>>& sets) {
    int total = 0;
    for (int k : sizes) total += k;
    if (total > n) return;
    std::vector<uint32_t> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < total; i++) {  // partial Fisher-Yates
        uint64_t j = i + r.below(n - i);
        std::swap(perm[i], perm[j]);
    }
    std::unordered_set<uint64_t> have;
    have.reserve(e.size() * 2 + 16);
    for (auto& p : e) have.insert(key(p.first, p.second));
    int at = 0;
    for (int k : sizes) {
        std::vector<uint32_t> mem(perm.begin() + at, perm.begin() + at + k);
        at += k;
        std::sort(mem.begin(), mem.end());
        sets.push_back(mem);
        // remove existing internal edges, then keep each pair w.p. 0.95
        std::vector<std::vector<char>> in(k, std::vector<char>(k, 0));
        for (int a = 0; a < k; a++)
            for (int b = a + 1; b < k; b++) {
                have.erase(key(mem[a], mem[b]));
                if (r.uni() < 0.95) in[a][b] = in[b][a] = 1;
            }
        int req = (int)std::ceil(g * (k - 1));
        std::vector<int> deg(k, 0);
        for (int a = 0; a < k; a++)
            for (int b = 0; b < k; b++) deg[a] += in[a][b];
        for (;;) {
            int lo = -1;
            for (int a = 0; a < k; a++)
                if (deg[a] < req && (lo < 0 || deg[a] < deg[lo])) lo = a;
            if (lo < 0) break;
            int pick = -1;
            for (int b = 0; b < k; b++)
                if (b != lo && !in[lo][b] && (pick < 0 || deg[b] < deg[pick])) pick = b;
            if (pick < 0) break;
            in[lo][pick] = in[pick][lo] = 1;
            deg[lo]++;
            deg[pick]++;
        }
        for (int a = 0; a < k; a++)
            for (int b = a + 1; b < k; b++)
                if (in[a][b]) have.insert(key(mem[a], mem[b]));
    }
    e.clear();
    e.reserve(have.size());
    for (uint64_t kk : have) e.push_back({(uint32_t)(kk >> 32), (uint32_t)(kk & 0xffffffffu)});
}

// |cos| threshold graph over rank-16 features drawn from `nmod` modules
// (Zipf sizes, member = centre + sigma z) and a share `noise` of isotropic
// vertices.
void cosine_edges(Rng& r, int n, int rank, int nmod, double sigma, double noise, uint64_t target, Edges& e) {
    std::vector<float> f((size_t)n * rank);
    std::vector<double> mu((size_t)nmod * rank);
    for (auto& x : mu) x = r.normal();
    std::vector<double> w(nmod);
    double sw = 0;
    for (int k = 0; k < nmod; k++) {
        w[k] = 1.0 / (k + 1.0);
        sw += w[k];
    }
    std::vector<int> lab(n, -1);
    int in_mod = (int)((1.0 - noise) * n), fill = 0;
    for (int k = 0; k < nmod; k++) {
        int sz = (int)(w[k] / sw * in_mod);
        for (int q = 0; q < sz && fill < n; q++) lab[fill++] = k;
    }
    for (int i = 0; i < n; i++) {
        double nn = 0;
        double tmp[64];
        for (int k = 0; k < rank; k++) {
            tmp[k] = lab[i] >= 0 ? mu[(size_t)lab[i] * rank + k] + sigma * r.normal() : r.normal();
            nn += tmp[k] * tmp[k];
        }
        nn = std::sqrt(nn);
        for (int k = 0; k < rank; k++) f[(size_t)i * rank + k] = (float)(tmp[k] / nn);
    }
    // all-pairs |cos| (float, fixed summation order), then the target-th
    // largest value is the threshold
    std::vector<float> c((size_t)n * (n - 1) / 2);
    size_t at = 0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            float s = 0;
            for (int k = 0; k < rank; k++) s += f[(size_t)i * rank + k] * f[(size_t)j * rank + k];
            c[at++] = std::fabs(s);
        }
    std::vector<float> tmp(c);
    size_t kth = tmp.size() - target;
    std::nth_element(tmp.begin(), tmp.begin() + kth, tmp.end());
    float thr = tmp[kth];
    at = 0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (c[at++] >= thr) e.push_back({(uint32_t)i, (uint32_t)j});
}

}  // namespace

int generate(int config, uint64_t seed, int instance, qc_graph* g, double* gamma, int* min_size) {
    Edges e;
    int n = 0;
    std::vector<std::vector<uint32_t>> planted;
    if (config == 0) {
        Rng r(seed);
        n = 200;
        er_edges(r, n, 0.08, e);
        plant(r, n, {20, 22, 25}, 0.9, e, planted);
        *gamma = 0.9;
        *min_size = 10;
    } else if (config == 1) {
        Rng r(seed);
        n = 8000;
        cosine_edges(r, n, 16, 32, 0.3, 0.1, 1000000, e);
        *gamma = 0.9;
        *min_size = 30;
    } else if (config == 2) {
        Rng r(seed);
        n = 100000;
        chung_lu_edges(r, n, 20.0, 2.1, e);
        std::vector<int> sizes(50);
        for (auto& k : sizes) k = 20 + (int)r.below(41);
        plant(r, n, sizes, 0.85, e, planted);
        *gamma = 0.85;
        *min_size = 20;
    } else if (config == 3) {
        Rng r(seed * 1000003ULL + (uint64_t)instance);
        bool er = r.uni() < 0.5;
        n = (int)std::llround(std::exp(std::log(500.0) + r.uni() * (std::log(20000.0) - std::log(500.0))));
        double d = 8.0 + r.uni() * 52.0;
        double gm = std::round((0.6 + r.uni() * 0.35) * 100.0) / 100.0;
        int ms = 8 + (int)r.below(33);
        if (er)
            er_edges(r, n, d / (n - 1), e);
        else
            chung_lu_edges(r, n, d, 2.1 + r.uni() * 0.7, e);
        int np = 1 + (int)r.below(10);
        std::vector<int> sizes(np);
        for (auto& k : sizes) k = 10 + (int)r.below(31);
        plant(r, n, sizes, gm, e, planted);
        *gamma = gm;
        *min_size = ms;
    } else if (config == 4) {
        Rng r(seed);
        n = 1000000;
        chung_lu_edges(r, n, 30.0, 2.2, e);
        std::vector<int> sizes(200);
        for (auto& k : sizes) k = 30 + (int)r.below(51);
        plant(r, n, sizes, 0.9, e, planted);
        *gamma = 0.9;
        *min_size = 40;
    } else {
        set_error("unknown generator config %d", config);
        return QC_EINVAL;
    }
    const int rc = csr_from_edges(n, e, g);
    if (rc == QC_OK) {
        g->planted_off.assign(1, 0);
        g->planted.clear();
        for (auto& m : planted) {
            g->planted.insert(g->planted.end(), m.begin(), m.end());
            g->planted_off.push_back(g->planted.size());
        }
    }
    return rc;
}

}  // namespace qc
