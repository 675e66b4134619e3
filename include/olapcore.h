/*
 * olapcore.h -- C ABI of the MI355X star-join group-by library (libolapcore.so).
 *
 * The reference (panjd123/memory_database) exposes its OLAPcore operator as
 * C++ free functions over raw column pointers, e.g.
 *   int OLAPcore_rowwise(int8_t **dimvec_array, int32_t **fk_array, int *orders,
 *                        int &dimvec_nums, int *factor, int &size_lineorder,
 *                        int &group_nums, int32_t *M1, int32_t *M2, int nthreads);
 * (code/OLAPcore/test_multi_compute_operator.cpp:973). This header carries the
 * same arguments -- per-table dimension vectors and foreign-key columns, the
 * probe order, the mixed-radix factors, the fact-row count, the group count
 * and the two measure columns -- in two plain structs, and returns explicit
 * status codes instead of exit(-1).
 *
 * Conventions
 *  - All column and output pointers are DEVICE pointers (HBM-resident) unless
 *    a parameter says "host". Streams are hipStream_t passed as void* (NULL =
 *    the null stream). Entry points that return a host value synchronise the
 *    stream; the others are asynchronous.
 *  - Dimension vectors hold one code per dimension row: int8 with NULL 127
 *    (DIM_NULL, include/metadata.h:69) or int32 with NULL -1.
 *  - A fact row passes when every probed dimension yields a non-NULL code. A
 *    foreign key outside [0, dim_rows) finds no dimension row and fails (the
 *    reference reads past the vector there, include/gendata_util.hpp:530-534).
 *  - gid = sum_i code_i * factor[i]; sum[gid] += (int64)m1 + (int64)m2 and
 *    count[gid] += 1, in int64 (the reference's uint32 vector wraps).
 *  - Results are bit-exact and independent of strategy and launch shape.
 *  - A context is not thread-safe; use one context per host thread.
 */
#ifndef OLAPCORE_H
#define OLAPCORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OLAP_MAX_DIMS 4
#define OLAP_DIM_NULL_I8 127    /* DIM_NULL, include/metadata.h:69 */
#define OLAP_DIM_NULL_I32 (-1)
#define OLAP_GID_NULL (-1)      /* widened GROUP_NULL (include/metadata.h:70) */
#define OLAP_STARJOIN_NULL 32767 /* GROUP_NULL as the star-join result uses it */

typedef enum {
    OLAP_OK = 0,
    OLAP_E_INVALID = 1,   /* bad argument or input data (message in olap_last_error) */
    OLAP_E_HIP = 2,       /* a HIP runtime call failed */
    OLAP_E_NOMEM = 3,     /* device allocation failed */
    OLAP_E_CAPACITY = 4,  /* request beyond a documented limit */
    OLAP_E_NODEVICE = 5,  /* no usable GPU */
    OLAP_E_BUSY = 6       /* admission refused (server) */
} olap_status;

typedef enum {
    OLAP_ROWWISE = 0, /* fused probe, per-row early exit   (OLAPcore_row_thread :329) */
    OLAP_VWM_DV = 1,  /* tile compaction, dynamic vector  (OLAPcore_vwm_dv_thread :375) */
    OLAP_VWM_SV = 2,  /* tile static vector, no early exit (OLAPcore_vwm_sv_thread :509) */
    OLAP_CWM_DV = 3,  /* column-wise, materialised lists  (join/agg_cwm_dv :240/:645) */
    OLAP_CWM_SV = 4   /* column-wise, full gid vector     (join/agg_cwm_sv :290/:662) */
} olap_strategy;

/* The plan main() builds, test_multi_compute_operator.cpp:1964-1987. */
typedef struct {
    int32_t dimvec_nums;              /* probe steps, 1..4 */
    int32_t orders[OLAP_MAX_DIMS];    /* table slot probed at step i */
    int32_t factor[OLAP_MAX_DIMS];    /* multiplier of step i's code, >= 0 */
    int64_t group_nums;               /* slots in the outputs, 1..2^31-1 */
} olap_plan;

/* The column arguments of the reference operators, indexed by table slot
 * (0 customer, 1 supplier, 2 part, 3 date as in the reference's main). */
typedef struct {
    const void *dimvec[OLAP_MAX_DIMS];  /* device; codes per dimension row */
    int64_t dim_rows[OLAP_MAX_DIMS];    /* entries in dimvec[t], < 2^31 */
    int32_t code_bytes[OLAP_MAX_DIMS];  /* 1 (int8, NULL 127) or 4 (int32, NULL -1) */
    const int32_t *fk[OLAP_MAX_DIMS];   /* device; foreign keys, rows entries */
    const int32_t *m1;                  /* device; measure */
    const int32_t *m2;                  /* device; second measure or NULL */
    int64_t rows;                       /* fact rows (size_lineorder), < 2^32 */
} olap_columns;

typedef struct olap_ctx_s *olap_ctx;
typedef struct olap_query_s *olap_query;

/* ---- context and memory ------------------------------------------------ */
olap_status olap_ctx_create(int device, olap_ctx *out);
olap_status olap_ctx_destroy(olap_ctx ctx);
const char *olap_last_error(olap_ctx ctx);
olap_status olap_device_info(olap_ctx ctx, int32_t *num_cus, int64_t *hbm_bytes,
                             char *arch_name, size_t arch_len);
olap_status olap_dev_alloc(olap_ctx ctx, size_t bytes, void **out);
olap_status olap_dev_free(olap_ctx ctx, void *p);
olap_status olap_copy_h2d(olap_ctx ctx, void *dst, const void *src, size_t bytes, void *stream);
olap_status olap_copy_d2h(olap_ctx ctx, void *dst, const void *src, size_t bytes, void *stream);
olap_status olap_stream_sync(olap_ctx ctx, void *stream);

/* ---- synthetic data (counter-based, reproducible on the CPU) ------------ */
/* out[i] = lo + floor(u * (hi - lo)) with u from SplitMix64(seed, stream, first_row + i) */
olap_status olap_gen_uniform_i32(olap_ctx ctx, int32_t *out, int64_t n, int64_t first_row,
                                 uint64_t seed, uint64_t stream_id, int32_t lo, int32_t hi,
                                 void *stream);
/* host: inverse-CDF thresholds of Zipf(theta) over d ranks (d entries) */
olap_status olap_zipf_table(int64_t d, double theta, uint64_t *thr_host);
/* out[i] = rank drawn through a device copy of the table above */
olap_status olap_gen_zipf_i32(olap_ctx ctx, int32_t *out, int64_t n, int64_t first_row,
                              uint64_t seed, uint64_t stream_id, const uint64_t *thr_dev,
                              int64_t d, void *stream);

/* ---- plan ---------------------------------------------------------------- */
/* sort_by_rate (:223-233, quirk kept) + group_nums/factor (:1964-1987).
 * groups[t] == 0 leaves table t out; bitmap[t] == 0 makes it filter-only. */
olap_status olap_make_plan(const int32_t groups[OLAP_MAX_DIMS], const int32_t bitmap[OLAP_MAX_DIMS],
                           const double sele[OLAP_MAX_DIMS], olap_plan *out);

/* ---- step 1: filter a dimension into a dimension vector ------------------- */
/* out[r] = (lo <= filter[r] < hi) ? (bitmap ? (group_mod > 0 ? group[r] % group_mod
 *          : group[r]) : 0) : NULL; code_bytes 1 or 4. The dimension-vector
 * build of code/TPCH_Q5/Q5.cpp:921-944 and include/gendata_util.hpp:473-527. */
olap_status olap_build_dimvec(olap_ctx ctx, const int32_t *filter, const int32_t *group,
                              int64_t rows, int32_t lo, int32_t hi, int32_t bitmap,
                              int32_t group_mod, int32_t code_bytes, void *out, void *stream);

/* General dimension predicates: a dimension vector from a program of up to
 * OLAP_MAX_TERMS terms over the dimension's columns, combined left to right
 * (acc = acc AND term, or acc OR term), as the reference chains its select
 * nodes (code/TPCH_Q5_operator/q5_operator.cpp:1288-1596, each node ANDed or
 * ORed with the one before) and builds snowflake dimension vectors through a
 * second table (code/TPCH_Q5/Q5.cpp:921-944, q5_operator.cpp:1472-1506):
 *   RANGE  lo <= col[r] < hi (the <, <=, >, >=, == and != forms of the
 *          reference's value selects, with negate for !=)
 *   IN     col[r] is one of values[0..nvalues) (host array, any order); a
 *          dictionary-coded string equality is IN over the string's code
 *   HOP    parent[col[r] + key_bias] is a non-NULL code of the parent
 *          dimension vector (int8 NULL 127 or int32 NULL -1; keys outside
 *          [0, parent_rows) find no row) -- one hop of a snowflake; the
 *          parent may itself be built by a program (two hops)
 * negate inverts a term. The row's code: OLAP_GROUP_NONE 0 (filter only);
 * OLAP_GROUP_COLUMN group_col[r] % group_mod (group_mod 0: the raw value);
 * OLAP_GROUP_HOP the parent code that term group_term found (a row whose
 * group term found none fails), e.g. customer -> nation grouped by nation. */
#define OLAP_MAX_TERMS 8
#define OLAP_MAX_IN_VALUES 4096
typedef enum { OLAP_TERM_RANGE = 0, OLAP_TERM_IN = 1, OLAP_TERM_HOP = 2 } olap_term_kind;
typedef enum { OLAP_AND = 0, OLAP_OR = 1 } olap_combine;
typedef enum { OLAP_GROUP_NONE = 0, OLAP_GROUP_COLUMN = 1, OLAP_GROUP_HOP = 2 } olap_group_kind;

typedef struct {
    int32_t kind;               /* olap_term_kind */
    int32_t combine;            /* olap_combine with the terms before; ignored for term 0 */
    int32_t negate;             /* 1: the term passes where its test fails */
    int32_t lo, hi;             /* RANGE */
    int32_t nvalues;            /* IN: 1..OLAP_MAX_IN_VALUES */
    int32_t parent_code_bytes;  /* HOP: 1 or 4 */
    int32_t key_bias;           /* HOP: added to col[r] (e.g. -1 for 1-based keys) */
    const int32_t *col;         /* device, dimension rows: the tested column or the hop key */
    const int32_t *values;      /* IN: host */
    const void *parent;         /* HOP: device parent dimension vector */
    int64_t parent_rows;        /* HOP */
} olap_term;

typedef struct {
    int32_t nterms;             /* 1..OLAP_MAX_TERMS */
    int32_t group;              /* olap_group_kind */
    int32_t group_mod;          /* OLAP_GROUP_COLUMN */
    int32_t group_term;         /* OLAP_GROUP_HOP: index of a HOP term, not negated */
    const int32_t *group_col;   /* OLAP_GROUP_COLUMN: device */
    olap_term term[OLAP_MAX_TERMS];
} olap_dim_program;

/* out[r] = program passes row r ? code : NULL, code_bytes 1 (codes must fit
 * 0..126, else OLAP_E_INVALID) or 4. Synchronises. */
olap_status olap_build_dimvec_program(olap_ctx ctx, int64_t rows, const olap_dim_program *prog,
                                      int32_t code_bytes, void *out, void *stream);

/* Dictionary coding of a string column (host): codes[i] = rank of strs[i]
 * among the column's distinct strings in byte order, so equal strings get
 * equal codes and the codes do not depend on row order or threads (the
 * reference's Group_Option numbers values first-seen under a mutex,
 * q5_operator.cpp:1639-1692). ndistinct (nullable) receives the count. */
olap_status olap_dict_encode(const char *const *strs, int64_t n, int32_t *codes, int64_t *ndistinct);
/* The code of `value` in a column coded by olap_dict_encode, -1 when absent. */
int32_t olap_dict_code_of(const char *const *strs, const int32_t *codes, int64_t n, const char *value);

/* ---- steps 2-4: probe, compose, group-by SUM ----------------------------- */
/* Validate and pack the dimension vectors once (synchronises). Rejects
 * negative non-NULL codes and plans whose largest gid reaches group_nums. */
olap_status olap_query_create(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                              olap_query *out);
/* Zero sum/count (group_nums int64 each, device) and run one strategy. */
olap_status olap_query_run(olap_query q, olap_strategy strategy, int64_t *sum, int64_t *count,
                           void *stream);
olap_status olap_query_destroy(olap_query q);
/* Launch shapes (threads per workgroup x rows per thread) of a strategy's
 * scan kernel, and the autotuner: it runs every shape of `strategy` on the
 * query (sum/count are overwritten, then hold a run of the winner), keeps the
 * fastest median-of-3 for this query and caches it in the context for
 * queries of the same class (strategy, probe steps, group-table residency,
 * log2 rows). Synchronises. No counterpart in the reference, which has one
 * fixed launch per strategy. */
int32_t olap_num_shapes(olap_strategy strategy);
olap_status olap_shape_info(olap_strategy strategy, int32_t shape, int32_t *threads,
                            int32_t *rows_per_thread);
/* How a row-wise shape fetches: 0 = one tile at a time; 1 = every dependent
 * fetch (the keys of the later probe steps, the measures) one tile ahead;
 * 2 = the same, never streaming the second step's keys with the first's
 * (1 streams them when most of their cache lines hold a live row); 3 / 4 =
 * survivor queue: the streamed keys two tiles ahead, the rows that pass them
 * queued per wave and finished 64 / 256 at a time by gathers; 5 / 6 = the
 * survivor queue never streaming the second step's keys. */
int32_t olap_shape_pipeline(olap_strategy strategy, int32_t shape);
/* Fix a row-wise query's launch shape (0..olap_num_shapes-1), or -1: the
 * shape the autotuner cached for the query's class, else a survivor-queue
 * shape (11 or 15, the latter when step 1's keys are better gathered than
 * streamed). */
olap_status olap_query_set_shape(olap_query q, int32_t shape);
/* Aggregation layout of row-wise runs whose group table exceeds LDS:
 * -1 automatic (from the expected matches: many matches take the one-pass
 * staged form, or the two-pass one where its open chunks would not fit);
 * 0 one global atomic pair per matching row; 1 two-pass staged (count,
 * scatter records to HBM, reduce per 8192-slot partition in LDS;
 * synchronises once per run to size the record buffer); 2 as 0 behind a
 * per-workgroup LDS cache of up to 4096 groups, which the first groups to
 * reach a slot keep (skewed keys: the hot groups stop contending on global
 * atomics); 3 one-pass staged (one scan writes every match's record into
 * chunks of a record pool, per partition; one reduce; no host
 * synchronisation; OLAP_E_CAPACITY where each workgroup's two open chunks
 * per partition would exceed 1/32 of HBM). Layouts 1 and 3 keep per-
 * partition state (8,192 groups a partition) in LDS, so both refuse group
 * tables above about 1.1e8 groups with OLAP_E_CAPACITY; the automatic
 * choice then takes global atomics (layout 0). The autotuner tries all four
 * and caches the winning shape and layout for queries of the same class,
 * whose automatic runs then use them (falling back to the automatic choice
 * where the cached staged layout does not fit). */
olap_status olap_query_set_aggregation(olap_query q, int32_t layout);
olap_status olap_query_autotune(olap_query q, olap_strategy strategy, int64_t *sum, int64_t *count,
                                void *stream, int32_t *shape_out, double *ms_out);
/* The launch of the query's last run (e.g. the autotuner's winner): threads
 * per workgroup, rows per thread and the aggregation used. OLAP_E_INVALID
 * before the first run. */
#define OLAP_AGG_ATOMICS 0   /* group table beyond LDS: global atomics */
#define OLAP_AGG_TWOPASS 1   /* two-pass staged */
#define OLAP_AGG_HOT 2       /* global atomics behind the LDS cache of hot groups */
#define OLAP_AGG_ONEPASS 3   /* one-pass staged */
#define OLAP_AGG_LDS 4       /* group table in LDS */
olap_status olap_query_launch_info(olap_query q, int32_t *threads, int32_t *rows_per_thread,
                                   int32_t *aggregation);
/* bytes of device memory the query holds (packed vectors, scratch) */
int64_t olap_query_device_bytes(olap_query q);
/* How the last row-wise run fetched: OLAP_PIPE_AHEAD when its dependent
 * fetches ran one tile ahead, OLAP_PIPE_QUEUE when it queued survivors,
 * OLAP_PIPE_STREAM2 when the second step's keys streamed with the first's;
 * -1 before any run. */
#define OLAP_PIPE_AHEAD 1
#define OLAP_PIPE_STREAM2 2
#define OLAP_PIPE_QUEUE 4
int32_t olap_query_pipeline(olap_query q);

/* One-shot forms named after the reference drivers; they create, run and
 * destroy a query and synchronise. total_count (host, nullable) receives the
 * sum of the count vector. */
olap_status olap_rowwise(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                         int64_t *sum, int64_t *count, int64_t *total_count, void *stream);
olap_status olap_vwm_dv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream);
olap_status olap_vwm_sv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream);
olap_status olap_cwm_dv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream);
olap_status olap_cwm_sv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream);

/* olap_rowwise for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Same arguments, results and errors as
 * olap_rowwise, but every piece of device scratch (dimension statistics, the
 * error word, the total count and the packed dimension vectors) lives in a
 * fixed per-GPU array of 8 MiB held until the process ends, and no device
 * memory is allocated or freed. Packed dimension vectors (1..32 bits per
 * code, rounded up per table to 16 bytes) beyond 8 MiB - 256 bytes are
 * refused with OLAP_E_CAPACITY. Group tables beyond LDS take global atomics.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several host threads or streams) race. Synchronises the stream. */
olap_status olap_rowwise_noalloc(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                                 int64_t *sum, int64_t *count, int64_t *total_count, void *stream);

/* The column-wise pairs split as in the reference: join_cwm_sv (:908) writes
 * a per-row gid vector (rows int32, device, OLAP_GID_NULL for failing rows),
 * agg_cwm_sv (:764) aggregates it. join_cwm_dv (:840) leaves the surviving
 * (row, gid) lists in the query; agg_cwm_dv (:685) aggregates them. */
olap_status olap_join_cwm_sv(olap_query q, int32_t *gid_vec, void *stream);
olap_status olap_agg_cwm_sv(olap_query q, const int32_t *gid_vec, int64_t *sum, int64_t *count,
                            void *stream);
olap_status olap_join_cwm_dv(olap_query q, void *stream);
olap_status olap_agg_cwm_dv(olap_query q, int64_t *sum, int64_t *count, void *stream);

/* host helper: the int every reference driver returns -- the uint32 group
 * vector summed into an int (e.g. :1027-1036) -- from int64 sums (host). */
int32_t olap_legacy_count(const int64_t *sum_host, int64_t group_nums);
/* The reference's raw result, its uint32 group vector (summed over threads,
 * :1022-1034): u32_out[g] = (uint32_t)sum[g] for g < group_nums (device
 * arrays). Where the reference's vector wraps, this wraps the same way;
 * asynchronous. */
olap_status olap_legacy_u32(olap_ctx ctx, const int64_t *sum, int64_t group_nums, uint32_t *u32_out,
                            void *stream);

/* ---- batched form: many star-join group-by queries in one scan ------------ */
/* No counterpart in the reference, which runs one query per scan. A batch
 * holds up to 64 queries over one shared star schema; a single persistent
 * kernel reads the fact table once and evaluates every query on every row.
 * Each query is given by its predicates and group-by expressions rather than
 * by pre-built dimension vectors: step i probes table[i], passes when
 * lo[i] <= filter[table[i]][key] < hi[i], and contributes the code
 * (uint32)attr[table[i]][key] % group_mod[i] (0 when group_mod[i] == 0, a
 * filter-only step) times factor[i] to the gid. The result of query q is the
 * one olap_query_run gives for the same plan over the int32 dimension vectors
 * olap_build_dimvec(filter, attr, lo, hi, group_mod > 0, group_mod, 4). */
#define OLAP_BATCH_MAX 64

typedef struct {
    int32_t ndims;                      /* probe steps, 1..4, distinct tables */
    int32_t table[OLAP_MAX_DIMS];       /* table slot of step i */
    int32_t lo[OLAP_MAX_DIMS];          /* predicate lo <= filter < hi */
    int32_t hi[OLAP_MAX_DIMS];
    int32_t group_mod[OLAP_MAX_DIMS];   /* code = attr % group_mod; 0: filter only */
    int32_t factor[OLAP_MAX_DIMS];      /* multiplier of step i's code */
    int64_t group_nums;                 /* output slots of the query */
    /* optional general predicate of step i (host, NULL: none), ANDed with
     * the range above; its group output is not used (the code is
     * attr % group_mod). Read at batch creation (a server reads it when the
     * request's batch is created: keep it valid until the wait returns). */
    const olap_dim_program *program[OLAP_MAX_DIMS];
} olap_batch_query;

/* The shared star schema: per table slot a filter and a group attribute
 * column (device, dim_rows[t] entries; NULL for tables no query probes), and
 * the fact columns. */
typedef struct {
    const int32_t *filter[OLAP_MAX_DIMS];
    const int32_t *attr[OLAP_MAX_DIMS];
    int64_t dim_rows[OLAP_MAX_DIMS];
    const int32_t *fk[OLAP_MAX_DIMS];
    const int32_t *m1;
    const int32_t *m2;                  /* nullable */
    int64_t rows;
} olap_star;

typedef struct olap_batch_s *olap_batch;

/* Validate the queries, build the per-table query masks (bit q of mask[t][r]
 * set when query q passes dimension row r of table t or does not probe t)
 * with each dimension row's group attribute beside its mask, and plan the
 * LDS layout (synchronises). The dimension columns are read here only: runs
 * use the masks and attributes as of creation (the fact columns are read by
 * every run). */
olap_status olap_batch_create(olap_ctx ctx, const olap_star *star, const olap_batch_query *queries,
                              int32_t nq, olap_batch *out);
/* offsets[q] = first output slot of query q, offsets[nq] = total slots (host) */
olap_status olap_batch_offsets(olap_batch b, int64_t *offsets);
/* Zero the concatenated outputs (total slots int64 each) and run the fused
 * scan. Asynchronous, no host synchronisation and no allocation: the call can
 * be captured into a hipGraph. */
olap_status olap_batch_run(olap_batch b, int64_t *sum, int64_t *count, void *stream);
/* Autotuner of the batched scan: times every (threads, rows per thread, LDS
 * budget for per-query group tables, aggregation of the other queries)
 * configuration on this batch, keeps the fastest; outputs then hold a run of
 * the winner. Synchronises. Aggregation (the `staged` output of
 * olap_batch_config_info): 0 global atomics; 1 two-pass staging (a count
 * pass sizes exact record ranges; the run reads the record total back once,
 * a captured graph reuses the buffers sized by the run made before capture;
 * matches beyond those buffers, after the columns have changed, are added
 * with global atomics); 2 one-pass staging in the event-grouped kernel
 * (records go to chunks of a pool sized from the expected matches, no host
 * read; matches beyond the pool are added with global atomics).
 * rows_per_thread 2 marks the event-grouped kernel. Configurations whose LDS
 * layout does not fit the batch (too many aggregation partitions), and
 * event-grouped configurations on fewer than 2 fact rows, are skipped. */
int32_t olap_batch_num_configs(void);
olap_status olap_batch_config_info(int32_t cfg, int32_t *threads, int32_t *rows_per_thread,
                                   int32_t *lds_table_kib, int32_t *staged);
/* The scan kernel a configuration runs: 0 per lane (every lane walks its
 * rows' matching queries), 1 event-grouped with counted event lists, 2
 * event-grouped with fixed per-query event segments; -1 for an invalid cfg. */
int32_t olap_batch_config_form(int32_t cfg);
olap_status olap_batch_set_config(olap_batch b, int32_t cfg);
olap_status olap_batch_autotune(olap_batch b, int64_t *sum, int64_t *count, void *stream,
                                int32_t *cfg_out, double *ms_out);
/* Capture zero + scan for these outputs into a hipGraph held by the batch,
 * then replay it; olap_batch_graph_launch is asynchronous. */
olap_status olap_batch_graph_capture(olap_batch b, int64_t *sum, int64_t *count, void *stream);
olap_status olap_batch_graph_launch(olap_batch b, void *stream);
/* Expected (row, query) matches per fact row, from the mask statistics
 * (uniform keys assumed); host, after create. */
double olap_batch_expected_events(olap_batch b);
olap_status olap_batch_destroy(olap_batch b);

/* ---- serving: continuous batching over resident datasets ----------------- */
/* No counterpart in the reference. A server keeps star schemas resident in
 * HBM and runs `streams` worker threads, each with its own context and
 * non-blocking HIP stream. An idle worker takes the oldest pending request,
 * admits further pending requests of the same dataset in arrival order while
 * their device bytes fit the free budget (at most max_batch), and runs them
 * as one fused batched scan; requests arriving meanwhile are admitted by the
 * next idle worker, so batches overlap on the device. A worker's results
 * come back into one of two pinned host buffers and a per-worker thread
 * copies each request's slice to its host buffers while the worker runs the
 * next batch. Results are exact: every request gets the same sums and counts
 * alone or in any mix. */
typedef struct {
    int32_t max_batch;       /* queries per fused scan, 1..64 */
    int32_t streams;         /* concurrent batches (worker threads), 1..8 */
    int64_t memory_budget;   /* device bytes admitted requests may hold */
} olap_server_config;

typedef struct {
    int64_t submitted, completed, failed, rejected, cancelled, batches;
    double mean_batch;        /* requests per fused scan */
    double p50_us, p99_us, max_us, mean_us;  /* submit-to-result latency */
    int64_t bytes_in_use, bytes_peak;
} olap_server_stats;

typedef struct olap_server_s *olap_server;

olap_status olap_server_create(olap_ctx ctx, const olap_server_config *cfg, olap_server *out);
/* A batch executor: runs queries[0..n) of one dataset into the concatenated
 * outputs sum/count (query i owns [offsets[i], offsets[i+1]), zeroed by the
 * executor). The library's own executor is the fused GPU scan (olap_batch);
 * a server made with olap_server_create_host calls `exec` instead, from its
 * worker threads, on host memory -- the seam that lets the admission,
 * batching, failure and cancellation logic run without a device. */
typedef olap_status (*olap_server_exec_fn)(void *user, int32_t dataset,
                                           const olap_batch_query *queries, int32_t n,
                                           const int64_t *offsets, int64_t *sum, int64_t *count);
olap_status olap_server_create_host(const olap_server_config *cfg, olap_server_exec_fn exec,
                                    void *exec_user, olap_server *out);
/* Output staging allocator (device memory for the GPU executor, host memory
 * for a host executor). A NULL return fails that batch's requests with
 * OLAP_E_NOMEM; the server keeps serving. Set before the first submit. */
typedef void *(*olap_server_alloc_fn)(void *user, size_t bytes);
typedef void (*olap_server_free_fn)(void *user, void *ptr);
olap_status olap_server_set_allocator(olap_server s, olap_server_alloc_fn alloc_fn,
                                      olap_server_free_fn free_fn, void *user);
/* register a resident star schema (pointers stay owned by the caller) */
olap_status olap_server_add_dataset(olap_server s, const olap_star *star, int32_t *dataset_id);
/* Queue one query. sum_host / count_host receive group_nums int64 each and
 * must stay valid until olap_server_wait returns for the ticket. Returns
 * OLAP_E_CAPACITY when the query alone exceeds the memory budget. */
olap_status olap_server_submit(olap_server s, int32_t dataset, const olap_batch_query *q,
                               int64_t *sum_host, int64_t *count_host, uint64_t *ticket);
/* Wait up to timeout_us (negative: forever) for a ticket; OLAP_E_BUSY on
 * timeout. On return *status holds the request's own status (OLAP_OK or the
 * error of its batch), *latency_us (optional) its submit-to-result time, and
 * the ticket is released. */
olap_status olap_server_wait(olap_server s, uint64_t ticket, int64_t timeout_us, int32_t *status,
                             double *latency_us);
/* Stop (paused = 1) or resume (0) admission; batches already admitted finish.
 * Requests submitted while paused are admitted together on resume. */
olap_status olap_server_pause(olap_server s, int32_t paused);
/* Withdraw a request that no worker has admitted yet (OLAP_E_BUSY otherwise). */
olap_status olap_server_cancel(olap_server s, uint64_t ticket);
olap_status olap_server_get_stats(olap_server s, olap_server_stats *out);
olap_status olap_server_reset_stats(olap_server s);
/* message of the last failure on the server; valid until the next failing
 * call on it */
const char *olap_server_last_error(olap_server s);
/* Stops the workers after their running batches (whose requests get their
 * results); requests not yet admitted are dropped, so call it once every
 * wait has returned. */
olap_status olap_server_destroy(olap_server s);
/* The admission accounting and policy, as pure host functions: the device
 * bytes a request holds (an int32 dimension vector per probed table plus its
 * int64 sum and count outputs), and the batch an idle worker takes from n
 * pending requests in arrival order -- the oldest request's dataset, arrival
 * order, at most max_batch, skipping requests that would overflow
 * budget_left; none when the oldest does not fit. Returns the number chosen
 * (indices into the pending list written to chosen). */
int64_t olap_server_request_bytes(const olap_star *star, const olap_batch_query *q);
int32_t olap_server_plan(const int32_t *datasets, const int64_t *bytes, int32_t n,
                         int64_t budget_left, int32_t max_batch, int32_t *chosen);

/* ---- sub-step operators --------------------------------------------------- */
/* starjoinalgo_rowwise / _cwm_sv / _cwm_dv (code/starjoin/test_starjoin_option.cpp:37-189):
 * rows passing all four int8 dimension vectors (c, s, p, d). result (device,
 * nullable) receives c<<12 | s<<8 | p<<4 | d per row, OLAP_STARJOIN_NULL when
 * the row fails. count is a host pointer. */
olap_status olap_starjoin(olap_ctx ctx, olap_strategy strategy, int64_t rows,
                          const int8_t *const dimvec[4], const int64_t dim_rows[4],
                          const int32_t *const fk[4], int32_t *result, int64_t *count,
                          void *stream);
/* vector_group (code/group/test_group_option.cpp:44-60): group_vector[key-1] +=
 * m1 + m2 over 1-based keys (int64, device, zeroed here); total (host) = sum
 * of the vector; bad_keys (host, nullable) = keys outside [1, group_num]. */
olap_status olap_vector_group(olap_ctx ctx, int64_t rows, const int32_t *key1,
                              const int32_t *m1, const int32_t *m2, int64_t *group_vector,
                              int64_t group_num, int64_t *total, int64_t *bad_keys,
                              void *stream);
/* vecjoin build (code/join/test_join_option.cpp:30-36): vector[key[i]] =
 * payload[i] over the build relation (device columns, rows < 2^31); a later
 * row overwrites an earlier row with the same key, as the reference's
 * sequential loop does; entries no key names keep their contents. Keys
 * outside [0, vector_len) are skipped and counted in bad_keys (host,
 * nullable). Synchronises. */
olap_status olap_join_build(olap_ctx ctx, int64_t rows, const int32_t *key, const int32_t *payload,
                            int32_t *vector, int64_t vector_len, int64_t *bad_keys, void *stream);
/* vecjoin probe (:58-64; probe_prefetch :72-82 sums the same): total (host)
 * = sum of vector[key[i]] over the probe relation, in int64 (the reference's
 * int count is its low 32 bits); keys outside [0, vector_len) add nothing
 * and are counted in bad_keys (host, nullable). A vector of up to 40,960
 * entries (160 KiB) is probed from LDS. Synchronises. */
olap_status olap_join_probe(olap_ctx ctx, int64_t rows, const int32_t *key, const int32_t *vector,
                            int64_t vector_len, int64_t *total, int64_t *bad_keys, void *stream);
/* TPC-H Q5's fused snowflake probe, Tri_table_join (code/TPCH_Q5/Q5.cpp:945-978,
 * driver :1000-1020): per fact (lineitem) row i, s = dimvec[0][key[0][i] +
 * key_bias[0]] (the supplier's nation code, NULL outside the region) and
 * o = dimvec[1][key[1][i] + key_bias[1]] (the ordering customer's nation code,
 * NULL outside the region or date range); the row counts when s is non-NULL
 * and s == o, and sum[s] += price[i] * (100 - discount[i]), count[s] += 1.
 * The reference sums l_extendedprice * (1 - l_discount) in double; here the
 * price is in cents and the discount in percent, so the int64 sum is exactly
 * 10^4 times the revenue. Outputs (device, group_nums int64 each) are zeroed
 * by the call; codes outside [0, group_nums) are refused (OLAP_E_INVALID).
 * Keys outside a vector find no row. Synchronises. */
#define OLAP_TRI_MAX_GROUPS 1024
typedef struct {
    int64_t rows;                 /* fact rows */
    const int32_t *key[2];        /* device: supplier key, order key per row */
    int32_t key_bias[2];          /* added to the keys (e.g. -1 for 1-based keys) */
    const void *dimvec[2];        /* device: supplier and order dimension vectors */
    int64_t dim_rows[2];
    int32_t code_bytes[2];        /* 1 (int8, NULL 127) or 4 (int32, NULL -1) */
    const int32_t *price;         /* device, cents */
    const int32_t *discount;      /* device, percent 0..100 */
    int32_t group_nums;           /* 1..OLAP_TRI_MAX_GROUPS */
} olap_tri_join;
olap_status olap_tri_table_join(olap_ctx ctx, const olap_tri_join *j, int64_t *sum, int64_t *count,
                                void *stream);
/* aggalgo_rowwise (code/aggregation/test_aggregation_option.cpp:28-37) in fixed
 * point: SUM(p*q*t + 100*p*q) with t the tax in percent; host result. */
olap_status olap_agg_expr(olap_ctx ctx, int64_t rows, const int32_t *p, const int32_t *q,
                          const int32_t *t, int64_t *result, void *stream);

#ifdef __cplusplus
}
#endif
#endif
