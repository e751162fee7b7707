// convert.cpp -- the reference's dataset converters as library calls:
//   edge list -> adjacency list      (edgeToAdj.py)
//   drop empty adjacency lines       (rmEmptyVert.py)
//   adjacency list -> serialized     (adjToSer.cpp:40-200)
// Each produces the bytes the reference's tool writes for the same input;
// inputs the reference cannot handle (a malformed edge line, an id that
// indexes past the vertex count) are errors here instead of crashes.

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "qc_host.h"

namespace {

using qc::set_error;

bool read_file(const char* path, std::string& text) {
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f.is_open()) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    text = ss.str();
    return true;
}

// Python's str.split() on whitespace followed by int(): false when a token
// is not an integer
bool split_ints(const std::string& line, std::vector<long long>& out) {
    out.clear();
    size_t i = 0, n = line.size();
    while (i < n) {
        while (i < n && isspace((unsigned char)line[i])) i++;
        if (i >= n) break;
        size_t j = i;
        while (j < n && !isspace((unsigned char)line[j])) j++;
        std::string tok = line.substr(i, j - i);
        // int() accepts an optional sign and underscores between digits
        std::string digits;
        size_t k = 0;
        bool neg = false;
        if (k < tok.size() && (tok[k] == '+' || tok[k] == '-')) neg = tok[k++] == '-';
        bool prev_digit = false;
        for (; k < tok.size(); k++) {
            if (isdigit((unsigned char)tok[k])) {
                digits += tok[k];
                prev_digit = true;
            } else if (tok[k] == '_' && prev_digit && k + 1 < tok.size() && isdigit((unsigned char)tok[k + 1])) {
                prev_digit = false;
            } else {
                return false;
            }
        }
        if (digits.empty()) return false;
        errno = 0;
        long long v = strtoll(digits.c_str(), nullptr, 10);
        if (errno == ERANGE) return false;
        out.push_back(neg ? -v : v);
        i = j;
    }
    return true;
}

// the lines Python's text-mode iteration yields (universal newlines: \n,
// \r\n and \r end a line), without their line ends
std::vector<std::string> py_lines(const std::string& text) {
    std::vector<std::string> lines;
    size_t s = 0;
    for (size_t i = 0; i < text.size(); i++) {
        if (text[i] == '\n' || text[i] == '\r') {
            lines.push_back(text.substr(s, i - s));
            if (text[i] == '\r' && i + 1 < text.size() && text[i + 1] == '\n') i++;
            s = i + 1;
        }
    }
    if (s < text.size()) lines.push_back(text.substr(s));
    return lines;
}

std::string join(const std::vector<long long>& v) {
    std::string s;
    for (size_t i = 0; i < v.size(); i++) {
        if (i) s += ' ';
        s += std::to_string(v[i]);
    }
    return s;
}

bool write_text(const char* path, const std::string& text) {
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
    return fclose(f) == 0 && ok;
}

}  // namespace

extern "C" {

// edgeToAdj.py: every line holds "u v"; u gets v appended (and v gets u with
// `duplicate`); vertices 0..max id are printed one list per line, neighbours
// in input order.
int qc_convert_edge_to_adj(const char* in_path, const char* out_path, int32_t duplicate) {
    if (!in_path || !out_path) return set_error("NULL path"), QC_EINVAL;
    std::string text;
    if (!read_file(in_path, text)) return set_error("cannot open %s", in_path), QC_EIO;
    std::map<long long, std::vector<long long>> adj;
    long long maxv = 0;
    std::vector<long long> tok;
    const std::vector<std::string> lines = py_lines(text);
    for (size_t ln = 0; ln < lines.size(); ln++) {
        if (!split_ints(lines[ln], tok) || tok.size() != 2) {
            set_error("%s:%zu: expected two integers", in_path, ln + 1);
            return QC_EIO;
        }
        const long long u = tok[0], v = tok[1];
        maxv = std::max(maxv, std::max(u, v));
        adj[u].push_back(v);
        if (duplicate) adj[v].push_back(u);
    }
    std::string out;
    for (long long x = 0; x <= maxv; x++) {
        auto it = adj.find(x);
        if (it != adj.end()) out += join(it->second);
        out += '\n';
    }
    if (!write_text(out_path, out)) return set_error("cannot write %s", out_path), QC_EIO;
    return QC_OK;
}

// rmEmptyVert.py: drop the empty lines and renumber the remaining vertices
// in line order; every value must name a non-empty line.
int qc_convert_rm_empty(const char* in_path, const char* out_path) {
    if (!in_path || !out_path) return set_error("NULL path"), QC_EINVAL;
    std::string text;
    if (!read_file(in_path, text)) return set_error("cannot open %s", in_path), QC_EIO;
    const std::vector<std::string> lines = py_lines(text);
    std::vector<std::vector<long long>> lists(lines.size());
    for (size_t ln = 0; ln < lines.size(); ln++)
        if (!split_ints(lines[ln], lists[ln])) return set_error("%s:%zu: not an integer", in_path, ln + 1), QC_EIO;
    std::vector<long long> map(lines.size(), -1);
    long long next = 0;
    for (size_t ln = 0; ln < lists.size(); ln++)
        if (!lists[ln].empty()) map[ln] = next++;
    std::string out;
    for (size_t ln = 0; ln < lists.size(); ln++) {
        if (lists[ln].empty()) continue;
        std::vector<long long> mapped;
        for (long long v : lists[ln]) {
            if (v < 0 || v >= (long long)map.size() || map[v] < 0)
                return set_error("%s:%zu: %lld is not a non-empty vertex", in_path, ln + 1, v), QC_EINVAL;
            mapped.push_back(map[v]);
        }
        out += join(mapped);
        out += '\n';
    }
    if (!write_text(out_path, out)) return set_error("cannot write %s", out_path), QC_EIO;
    return QC_OK;
}

// adjToSer.cpp: parse the adjacency text character by character exactly as
// adjToSer.cpp:60-110 does (a line break closes a vertex, so a final line
// break adds an empty last vertex; a space always closes a number), build
// N2(v) = N(v) u N(N(v)) \ {v} (adjToSer.cpp:115-156; a self loop stays in
// the 1-hop part), sort both lists and write the seven records.
int qc_convert_adj_to_ser(const char* in_path, const char* out_path) {
    if (!in_path || !out_path) return set_error("NULL path"), QC_EINVAL;
    std::string text;
    if (!read_file(in_path, text)) return set_error("invalid graph file"), QC_EIO;
    std::vector<int32_t> nb;
    std::vector<uint64_t> off{0};
    uint32_t cur = 0;
    bool empty = true;
    for (char ch : text) {
        if (ch == '\n') {
            if (!empty) nb.push_back((int32_t)cur);
            off.push_back(nb.size());
            cur = 0;
            empty = true;
        } else if (ch == ' ') {
            nb.push_back((int32_t)cur);
            cur = 0;
        } else if (ch == '\0') {
            break;
        } else {
            cur = cur * 10u + (uint32_t)(int32_t)(ch - '0');
            empty = false;
        }
    }
    if (!empty) nb.push_back((int32_t)cur);
    off.push_back(nb.size());
    const int64_t V = (int64_t)off.size() - 1;
    for (int32_t x : nb)
        if (x < 0 || x >= V) return set_error("neighbour %d outside 0..%lld", x, (long long)V - 1), QC_EINVAL;
    std::vector<int32_t> n2;
    std::vector<uint64_t> off2{0};
    std::vector<char> flag(V, 1);
    for (int64_t i = 0; i < V; i++) {
        const uint64_t s2 = n2.size();
        for (uint64_t j = off[i]; j < off[i + 1]; j++) {
            const int32_t a = nb[j];
            if (flag[a]) {
                n2.push_back(a);
                flag[a] = 0;
            }
            for (uint64_t k = off[a]; k < off[a + 1]; k++) {
                const int32_t b = nb[k];
                if (flag[b] && b != i) {
                    n2.push_back(b);
                    flag[b] = 0;
                }
            }
        }
        off2.push_back(n2.size());
        for (uint64_t j = s2; j < n2.size(); j++) flag[n2[j]] = 1;
        std::sort(nb.begin() + off[i], nb.begin() + off[i + 1]);
        std::sort(n2.begin() + s2, n2.end());
    }
    std::string out;
    auto put_list = [&](auto& v) {
        for (size_t i = 0; i < v.size(); i++) {
            if (i) out += ' ';
            out += std::to_string(v[i]);
        }
        out += '\n';
    };
    out += std::to_string(V) + "\n" + std::to_string(nb.size()) + "\n" + std::to_string(n2.size()) + "\n";
    put_list(nb);
    put_list(off);
    put_list(n2);
    put_list(off2);
    if (!write_text(out_path, out)) return set_error("cannot write %s", out_path), QC_EIO;
    return QC_OK;
}

}  // extern "C"
