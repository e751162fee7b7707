/*
 * qcmine.h -- C ABI of the MI355X (gfx950) maximal gamma-quasi-clique miner.
 *
 * The reference (Mike12041204/cuQC) has no library interface: its boundary is
 * the command line `cuQC graph_file gamma min_size output_file
 * scheduling_toggle` (main.cu:416-519), the serialized graph file
 * (main.cu:96-122, written by adjToSer.cpp:159-200) and the output file
 * format (main.cu:4641-4651).  Each entry point below names the reference
 * routine or behaviour it replaces.  The `qcmine` CLI (cuqc_amd/csrc/
 * qcmine_main.cpp) is built on these calls and keeps the reference's
 * arguments, messages and return codes.
 *
 * Conventions: every call returns a qc_status; on failure
 * qc_last_error() returns a thread-local message.  The library never exits
 * the process.  Handles are owned by the library and released with the
 * matching *_free call.  Plain pointers and sizes only.
 */
#ifndef QCMINE_H
#define QCMINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum qc_status {
    QC_OK = 0,
    QC_EINVAL = 1,    /* bad argument (gamma outside [0.5,1], min_size <= 1, ...) */
    QC_EIO = 2,       /* unreadable, short or corrupt file */
    QC_ENOMEM = 3,    /* host or device allocation failed */
    QC_EOVERFLOW = 4, /* a configured capacity was exceeded; no partial output */
    QC_EDEVICE = 5,   /* HIP runtime error or no usable gfx950 device */
    QC_EEMPTY = 6     /* zero quasi-cliques found (main.cu:4833, README.md:108) */
} qc_status;

typedef struct qc_graph qc_graph;
typedef struct qc_result qc_result;

/* Execution form of the enumeration. */
typedef enum qc_mode {
    QC_MODE_PLAIN = 0,      /* level-synchronous kernel, host loop per level
                               (the structure of search(), main.cu:606-678) */
    QC_MODE_PERSISTENT = 1  /* one persistent launch, device task queues */
} qc_mode;

/* Task order inside the persistent kernel. */
typedef enum qc_order {
    QC_ORDER_FIFO = 0,      /* plain LIFO task stack */
    QC_ORDER_COST = 1       /* MLP cost model orders work; output unchanged */
} qc_order;

typedef struct qc_params {
    double gamma;            /* main.cu:440-444, in [0.5, 1] */
    int32_t min_size;        /* main.cu:445-449, > 1 */
    int32_t scheduling;      /* main.cu:450-453: 0 dynamic, else static */
    int32_t mode;            /* qc_mode */
    int32_t order;           /* qc_order */
    int32_t workgroups;      /* 0 = auto (persistent: CUs x occupancy) */
    int32_t lds_tile;        /* largest task (vertices) staged in LDS; 0 = auto */
    int32_t split_chunk;     /* persistent mode: branches per work item when a task above
                                split_min is split across waves; 0 = 64, < 0 = never split */
    int32_t max_depth;       /* search window: expand tasks of depth <= max_depth; <= 0: all */
    uint64_t task_capacity;  /* vertex slots per task pool; 0 = auto */
    uint64_t clique_capacity;/* raw-clique vertex slots; 0 = auto */
    const void* cost_model;  /* qc_model*, required when order == QC_ORDER_COST */
    int64_t root_begin;      /* search window: root branches [root_begin, root_end) in the */
    int64_t root_end;        /*   reference's branch order (main.cu:1094); end <= 0: to the last;
                                begin < 0: that many branches before the last one's end */
    int32_t replays;         /* persistent mode: launch the captured hipGraph this many times and
                                require identical counters from every replay; <= 1: once */
    int32_t trace;           /* persistent mode: record the cost-model features of every expanded
                                task and its subtree size (qc_result_trace) */
    int32_t split_min;       /* persistent mode: split only tasks above this many vertices (and
                                above the LDS tile); 0 = the LDS tile for a single graph, 512 for
                                a batch */
    int32_t mid_tile;        /* persistent mode: tasks above the LDS tile and up to this size
                                (<= 4096) build task-local bit matrices in HBM; <= 0 = none (read
                                the root matrices) */
    int32_t lds_wide;        /* persistent mode: tasks above the LDS tile and up to this size
                                (<= 1024) keep their per-vertex arrays in LDS; 0 = 512, < 0 = none */
} qc_params;

typedef struct qc_stats {
    uint64_t tasks;          /* task records expanded, root included */
    uint64_t branches;       /* add-one-vertex steps */
    uint64_t children;       /* child task records written */
    uint64_t raw_cliques;    /* raw sets emitted before RemoveNonMax */
    uint64_t max_cliques;    /* maximal sets after RemoveNonMax */
    uint64_t sum_t;          /* sum of expanded task sizes (vertices) */
    uint64_t sum_t_written;  /* sum of written child sizes */
    uint64_t sum_clique_len; /* sum of raw set sizes */
    uint64_t root_size;      /* vertices in the root task */
    uint64_t levels;         /* plain mode: host loop iterations */
    uint64_t large_tasks;    /* tasks above the LDS tile (HBM-resident) */
    uint64_t beyond;         /* child records past the depth window (not expanded) */
    uint64_t root_branches;  /* branches of the root task (window range) */
    uint64_t w_adj;          /* adjacency entries the reference's list scans visit (root-restricted) */
    uint64_t pool_peak;      /* persistent mode: high-water mark of the task pool (vertex slots,
                                root records included); records are recycled, so this is the
                                search frontier's size, not the sum of all records written */
    uint64_t reserved;
    double t_init_ms;        /* root construction + device tables */
    double t_enum_ms;        /* ENUMERATION window (main.cu:551-685) */
    double t_nonmax_ms;      /* RemoveNonMax */
    double t_kernel_ms;      /* expansion kernels only (HIP events) */
} qc_stats;

/* ---- graphs ---------------------------------------------------------- */

/* Replaces CPU_Graph(ifstream&) (main.cu:96-122): reads the 7-record
 * serialized text file (V, E, L2, N, N offsets, N2, N2 offsets). */
int qc_graph_load_serialized(const char* path, qc_graph** out);

/* Replaces adjToSer.cpp:74-107: adjacency-list text, line i = neighbours of
 * vertex i.  N2 is not built here (the miner derives it). */
int qc_graph_load_adjlist(const char* path, qc_graph** out);

/* Replaces adjToSer.cpp:115-200: writes the serialized text format with the
 * full-graph N2 lists (sorted, self excluded). */
int qc_graph_write_serialized(const qc_graph* g, const char* path);

/* Replaces adjToSer.cpp:115-200 for large graphs: the same file as
 * qc_graph_write_serialized, byte for byte, with the N2 lists counted and
 * built on the GPU (a visited bitmap per vertex in LDS) and written in chunks,
 * so host memory holds one chunk of N2 entries.  n is limited by the LDS a
 * workgroup may hold (about 1.31 million vertices on gfx950; QC_EINVAL
 * above); QC_EDEVICE without a GPU. */
int qc_graph_write_serialized_device(const qc_graph* g, const char* path);

/* Binary CSR file (header "QCG1", n, nnz, offsets u64, adjacency u32); the
 * format for graphs whose text / N2 form is impractical. */
int qc_graph_load_binary(const char* path, qc_graph** out);
int qc_graph_write_binary(const qc_graph* g, const char* path);

/* Builds a graph from symmetric CSR (copied; lists are sorted, self loops and
 * duplicates rejected with QC_EINVAL). */
int qc_graph_from_csr(int32_t n, const uint64_t* offsets, const uint32_t* adj, qc_graph** out);

/* Seeded synthetic workload `config` (0..4); config 3 takes `instance` in
 * [0, 256).  Also returns the workload's gamma and min_size. */
int qc_graph_generate(int32_t config, uint64_t seed, int32_t instance, qc_graph** out, double* gamma,
                      int32_t* min_size);

/* The quasi-cliques a generated graph was planted with (no reference
 * counterpart: ground truth for tests), `count` ascending vertex sets;
 * zero sets for graphs that were not generated. */
int qc_graph_planted(const qc_graph* g, uint64_t* count, const uint64_t** offsets, const uint32_t** verts);

/* Read-only views; pointers stay valid until qc_graph_free. */
int qc_graph_csr(const qc_graph* g, int32_t* n, const uint64_t** offsets, const uint32_t** adj);
uint64_t qc_graph_checksum(const qc_graph* g); /* FNV-1a over offsets and adjacency */
void qc_graph_free(qc_graph* g);

/* ---- mining ---------------------------------------------------------- */

/* Replaces search() + RemoveNonMax() (main.cu:535-692, 4939-4961): builds
 * the root task, enumerates on the GPU to completion, collects raw sets and
 * removes non-maximal ones on the device.  QC_EEMPTY when nothing is found
 * (the result is still returned, with zero sets). */
int qc_mine(const qc_graph* g, const qc_params* p, qc_result** out);

/* Many independent instances through one persistent launch (no reference
 * counterpart; each instance's result equals its own qc_mine). */
int qc_mine_batch(const qc_graph* const* graphs, const qc_params* params, int32_t count, qc_result** out);

/* Maximal sets, each sorted ascending, in lexicographic order
 * (OutputMaxSet's trie order, main.cu:4881-4937). */
int qc_result_view(const qc_result* r, uint64_t* count, const uint64_t** offsets, const uint32_t** verts);

/* Raw sets as emitted (vertex order of the emitting step; set order is
 * unspecified). */
int qc_result_raw(const qc_result* r, uint64_t* count, const uint64_t** offsets, const uint32_t** verts);

int qc_result_stats(const qc_result* r, qc_stats* s);

/* OutputOneSet format (main.cu:4641-4651): "%d " for the length and each id,
 * then "\n".  QC_EEMPTY (and no file) when the result holds no sets. */
int qc_result_write(const qc_result* r, const char* path);

/* Task trace of a run with qc_params.trace set (no reference counterpart):
 * `rows` expanded tasks, 12 features each (members, candidates, min/max/mean
 * indeg, min/max/mean exdeg, depth, ub - lb, mean |N2 n root| over
 * candidates, size), the row of the parent task (-1 for the root) and the
 * number of tasks in the subtree rooted at the task. */
int qc_result_trace(const qc_result* r, uint64_t* rows, const float** feats, const int64_t** parent,
                    const int64_t** subtree);

/* temp.txt-compatible raw dump (flush_cliques, main.cu:1234-1246). */
int qc_result_write_raw(const qc_result* r, const char* path);

void qc_result_free(qc_result* r);

/* ---- RemoveNonMax on its own ----------------------------------------- */

/* Replaces RemoveNonMax(temp, out) (main.cu:4939-4961) for caller-provided
 * raw sets: device canonicalisation, de-duplication and subset removal. */
int qc_remove_nonmax(uint64_t count, const uint64_t* offsets, const uint32_t* verts, qc_result** out);

/* RemoveNonMax over raw sets already in device memory, with no device
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  Every buffer it needs lives in a 64 MiB device array that the
 * library holds on each GPU it runs on until the process ends; inputs that
 * need more (about 64 B per set + 20 B per raw vertex + 12 B per vertex id
 * below n_vertices) fail with QC_EOVERFLOW before anything is launched.
 *   d_offsets[count + 1], d_verts[n_verts]: the raw sets (device memory),
 *     n_verts = d_offsets[count]; every vertex id is below n_vertices.
 *   d_out_count: receives the number of maximal sets, or UINT64_MAX when a
 *     vertex id is not below n_vertices.
 *   d_out_offsets[count + 1], d_out_verts[n_verts]: the maximal sets, each
 *     ascending, in lexicographic order (as qc_remove_nonmax).
 * The work is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns without waiting: outputs are valid once the stream
 * has reached this point.  It makes no hipMalloc, hipFree, memset or
 * synchronizing call.  All calls on one GPU share the scratch array, which
 * keeps the last call's contents: calls whose work overlaps on one GPU (other
 * streams, other host threads) race and must be ordered by the caller. */
int qc_remove_nonmax_noalloc(uint64_t count, uint64_t n_verts, const uint64_t* d_offsets, const uint32_t* d_verts,
                             uint32_t n_vertices, uint64_t* d_out_count, uint64_t* d_out_offsets,
                             uint32_t* d_out_verts, void* stream);

/* ---- task-cost model (no reference counterpart) ------------------------
 * A 12 -> 64 -> 64 -> 1 ReLU MLP with bfloat16 weights and fp32 accumulation
 * that predicts log(subtree size) of a task from its qc_result_trace
 * features; trained on the GPU.  Passed as qc_params.cost_model with
 * order = QC_ORDER_COST it orders the persistent kernel's work; mined sets
 * never depend on it. */
typedef struct qc_model qc_model;
#define QC_MODEL_NPARAMS 5057   /* W1[12][64] b1[64] W2[64][64] b2[64] W3[64] b3 */
#define QC_MODEL_NNORM 29       /* mu[12] sd[12] ymu ysd thresholds[3] */

/* Adam on the mean squared error of standardized log targets, full batch,
 * `steps` updates; norms are fitted to the data. */
int qc_model_train(const float* feats, const float* log_target, uint64_t n, int32_t steps, float lr, uint64_t seed,
                   qc_model** out);
/* Predicted log(subtree size) for n feature rows. */
int qc_model_predict(const qc_model* m, const float* feats, uint64_t n, float* out);
/* Gradient of the loss (mean squared error on standardized targets) with
 * respect to the QC_MODEL_NPARAMS parameters, computed by the GPU kernels. */
int qc_model_grad(const qc_model* m, const float* feats, const float* log_target, uint64_t n, float* grad,
                  float* loss);
/* Untrained model: seeded He initialisation, identity norms. */
int qc_model_create(uint64_t seed, qc_model** out);
/* Read / replace fp32 parameters and the QC_MODEL_NNORM normalisation values
 * (any pointer may be NULL); the bf16 copy follows the fp32 values. */
int qc_model_params(qc_model* m, float* get, const float* set, float* norm_get, const float* norm_set);
int qc_model_save(const qc_model* m, const char* path);
int qc_model_load(const char* path, qc_model** out);
void qc_model_free(qc_model* m);

/* ---- tables ---------------------------------------------------------- */

/* Replaces calculate_minimum_degrees (main.cu:526-533) and the floor(d/gamma)
 * of the LU bounds (main.cu:1905): mindeg[i] = ceil(gamma*(i-1)) for
 * i < n_mindeg (mindeg[0] = 0) and ubq[d] = floor(d/gamma) for d < n_ubq, in
 * the reference's double arithmetic. */
int qc_degree_tables(double gamma, int32_t n_mindeg, int32_t* mindeg, int32_t n_ubq, int32_t* ubq);

/* ---- dataset converters (the reference's preprocessing tools) ---- */

/* edgeToAdj.py: "u v" lines -> one adjacency line per vertex 0..max id,
 * neighbours in input order; duplicate != 0 also appends u to v's line. */
int qc_convert_edge_to_adj(const char* in_path, const char* out_path, int32_t duplicate);

/* rmEmptyVert.py: drop empty adjacency lines and renumber the others in line
 * order (a value naming an empty line is QC_EINVAL). */
int qc_convert_rm_empty(const char* in_path, const char* out_path);

/* adjToSer.cpp (main, CPU_Graph, write_serialized): adjacency text -> the
 * 7-record serialized graph with N2 lists, byte for byte. */
int qc_convert_adj_to_ser(const char* in_path, const char* out_path);

/* ---- misc ------------------------------------------------------------ */

const char* qc_last_error(void);
const char* qc_version(void);
/* 1 if a gfx950 device is usable, else 0 (no error is raised). */
int qc_device_available(void);

#ifdef __cplusplus
}
#endif

#endif /* QCMINE_H */
