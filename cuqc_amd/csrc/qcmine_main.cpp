// qcmine_main.cpp -- command line with the reference's surface (main.cu:416-519):
//
//   qcmine <graph_file> <gamma> <min_size> <output_file> <scheduling_toggle>
//
// Same argument checks, messages, return codes and ">:" / "--->:" stdout
// markers.  graph_file is the serialized text format (main.cu:96-122) or a
// binary CSR file ("QCG1").  Environment knobs (not part of the reference):
//   QCMINE_MODE=plain|persistent   execution form (default persistent)
//   QCMINE_RAW=<path>              also write the raw sets (temp.txt format)

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "../../include/qcmine.h"

static double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

int main(int argc, char* argv[]) {
    auto start_total = std::chrono::steady_clock::now();
    if (argc != 6) {
        printf("Usage: ./main <graph_file> <gamma> <min_size> <output_file.txt> <scheduling toggle 0-dyanmic/1-static>\n");
        return 1;
    }
    FILE* probe = fopen(argv[1], "rb");
    if (!probe) {
        printf("invalid graph file\n");
        return 1;
    }
    char magic[4] = {0, 0, 0, 0};
    size_t got = fread(magic, 1, 4, probe);
    fclose(probe);
    double gamma = atof(argv[2]);
    if (gamma < .5 || gamma > 1) {
        printf("minimum degree ratio must be between .5 and 1 inclusive\n");
        return 1;
    }
    int min_size = atoi(argv[3]);
    if (min_size <= 1) {
        printf("minimum size must be greater than 1\n");
        return 1;
    }
    int toggle = atoi(argv[5]);
    if (!(toggle == 0 || toggle == 1)) std::cout << "scheduling toggle must be 0 or 1" << std::endl;

    auto start = std::chrono::steady_clock::now();
    std::cout << ">:PRE-PROCESSING" << std::endl;
    qc_graph* g = nullptr;
    int rc = (got == 4 && memcmp(magic, "QCG1", 4) == 0) ? qc_graph_load_binary(argv[1], &g)
                                                          : qc_graph_load_serialized(argv[1], &g);
    if (rc != QC_OK) {
        printf("%s\n", qc_last_error());
        return 1;
    }
    std::cout << "--->:LOADING TIME: " << (long long)ms_since(start) << " ms" << std::endl;

    qc_params p;
    memset(&p, 0, sizeof(p));
    p.gamma = gamma;
    p.min_size = min_size;
    p.scheduling = toggle;
    const char* mode = getenv("QCMINE_MODE");
    p.mode = (mode && strcmp(mode, "plain") == 0) ? QC_MODE_PLAIN : QC_MODE_PERSISTENT;
    p.order = QC_ORDER_FIFO;

    std::cout << ">:INITIALIZING TASKS" << std::endl;
    std::cout << ">:BEGINNING EXPANSION" << std::endl;
    qc_result* r = nullptr;
    rc = qc_mine(g, &p, &r);
    if (rc != QC_OK && rc != QC_EEMPTY) {
        printf("%s\n", qc_last_error());
        qc_graph_free(g);
        return 1;
    }
    qc_stats st;
    qc_result_stats(r, &st);
    std::cout << "--->:ENUMERATION TIME: " << (long long)st.t_enum_ms << " ms" << std::endl;
    const char* raw = getenv("QCMINE_RAW");
    if (raw) qc_result_write_raw(r, raw);
    std::cout << ">:REMOVING NON-MAXIMAL CLIQUES" << std::endl;
    if (rc == QC_EEMPTY) {
        printf("Error: no quasi-cliques found\n");
        qc_result_free(r);
        qc_graph_free(g);
        return 1;
    }
    if (qc_result_write(r, argv[4]) != QC_OK) {
        printf("%s\n", qc_last_error());
        qc_result_free(r);
        qc_graph_free(g);
        return 1;
    }
    printf(">:NUMBER OF MAXIMAL CLIQUES: %d\n", (int)st.max_cliques);
    fflush(stdout);
    std::cout << "--->:REMOVE NON-MAX TIME: " << (long long)st.t_nonmax_ms << " ms" << std::endl;
    std::cout << "--->:TOTAL TIME: " << (long long)ms_since(start_total) << " ms" << std::endl;
    std::cout << ">:PROGRAM END" << std::endl;
    qc_result_free(r);
    qc_graph_free(g);
    return 0;
}
