// qc_host.h -- host-side types shared by the qcmine library sources.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qcmine.h"

// Undirected graph in CSR form (symmetric adjacency, sorted lists, no self
// loops).  Optional N2 lists are kept when the graph came from a serialized
// file, which carries them (main.cu:96-122); otherwise the miner derives the
// 2-hop relation itself.
struct qc_graph {
    int32_t n = 0;
    std::vector<uint64_t> off;   // n + 1
    std::vector<uint32_t> adj;   // off[n]
    // generated graphs: the vertex sets of the planted quasi-cliques
    std::vector<uint64_t> planted_off{0};
    std::vector<uint32_t> planted;
    bool has_n2 = false;
    std::vector<uint64_t> off2;  // n + 1 when has_n2
    std::vector<uint32_t> adj2;
    // device-resident copy of the arrays above, made by the first mining call
    // and reused by later ones (owned; released with the graph)
    void* dev = nullptr;
    void (*dev_free)(void*) = nullptr;
    qc_graph() = default;
    qc_graph(const qc_graph&) = delete;
    qc_graph& operator=(const qc_graph&) = delete;
    ~qc_graph() {
        if (dev && dev_free) dev_free(dev);
    }
};

namespace qc {

// thread-local error text behind qc_last_error()
void set_error(const char* fmt, ...);

// graph helpers (graph_io.cpp)
int csr_from_edges(int n, std::vector<std::pair<uint32_t, uint32_t>>& edges, qc_graph* g);
uint64_t fnv1a_graph(const qc_graph* g);

// generators (gen.cpp)
int generate(int config, uint64_t seed, int instance, qc_graph* g, double* gamma, int* min_size);

// integer tables replacing the reference's two double expressions
// (main.cu:531 and main.cu:1905): mindeg[i] = ceil(gamma (i-1)),
// ubq[d] = floor(d / gamma).
void build_tables(double gamma, int n, int min_size, std::vector<int32_t>& mindeg, std::vector<int32_t>& ubq);

}  // namespace qc
