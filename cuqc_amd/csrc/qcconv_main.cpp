// qcconv -- the reference's dataset preprocessing tools as one command:
//
//   qcconv edge-to-adj <edge_file> <0|1>      adjacency list on stdout (edgeToAdj.py)
//   qcconv rm-empty <adj_file> <out_file>     drop empty vertices (rmEmptyVert.py)
//   qcconv adj-to-ser <adj_file> <out_file>   serialized graph (adjToSer.cpp)

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

extern "C" {
#include "../../include/qcmine.h"
}

static int usage() {
    printf("Usage: qcconv edge-to-adj <input_file> <0 - normal edges / 1 - duplicate edges> >output.txt\n"
           "       qcconv rm-empty <input_file> <output_file>\n"
           "       qcconv adj-to-ser <graph_file> <output_file>\n");
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 4) return usage();
    const std::string cmd = argv[1];
    int rc;
    if (cmd == "edge-to-adj") {
        rc = qc_convert_edge_to_adj(argv[2], "/dev/stdout", atoi(argv[3]) == 1);
    } else if (cmd == "rm-empty") {
        rc = qc_convert_rm_empty(argv[2], argv[3]);
    } else if (cmd == "adj-to-ser") {
        rc = qc_convert_adj_to_ser(argv[2], argv[3]);
        if (rc == QC_EIO && strcmp(qc_last_error(), "invalid graph file") == 0) {
            printf("invalid graph file\n");
            return 1;
        }
    } else {
        return usage();
    }
    if (rc != QC_OK) {
        fprintf(stderr, "qcconv: %s\n", qc_last_error());
        return 1;
    }
    return 0;
}
