// olap_internal.h -- shared host/device definitions of libolapcore.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "olapcore.h"
#include "olap_stage.h"

namespace olap {

constexpr int kBlock = 512;              // threads per workgroup of the scan kernels
constexpr int kChunk = 4;                // rows per 128-bit load
constexpr int kUnroll = 4;               // 128-bit chunks per thread per tile (row-wise, vwm_sv)
constexpr int kTileRW = kBlock * kChunk * kUnroll;  // 8192 rows
constexpr int kTileV = 4096;             // rows per tile of the compacting kernels
constexpr int kPerThreadV = kTileV / kBlock;        // list entries per thread (8)
constexpr uint32_t kLdsBytes = 160 * 1024;          // LDS per CU on gfx950
constexpr uint32_t kHotEmpty = 0xffffffffu;         // an unclaimed slot of the LDS cache of hot groups
constexpr uint32_t kLdsListBytes = kTileV * 2 + kTileV * 4 + 16;  // u16 off + u32 gid + counter
// device scratch of the allocation-free row-wise entry point, per GPU:
// statistics, error word and total in the first kNoallocHeader bytes, the
// packed dimension vectors after them
constexpr size_t kNoallocScratchBytes = size_t(8) << 20;
constexpr size_t kNoallocHeader = 256;

// A dimension program as the kernel reads it (olap_dim_program with IN
// values in a sorted device copy).
struct DevTerm {
    int32_t kind, combine, negate, lo, hi, nvalues, parent_code_bytes, key_bias;
    const int32_t *col;
    const int32_t *values;
    const void *parent;
    int64_t parent_rows;
};
struct DevProgram {
    int32_t nterms, group, group_mod, group_term;
    const int32_t *group_col;
    DevTerm term[OLAP_MAX_TERMS];
};

struct ScanShape {
    int bs;      // threads per workgroup
    int unroll;  // 128-bit chunks per thread per tile
    int pipe;    // row-wise: 0 one tile at a time, 1 / 2 dependent fetches one tile
                 // ahead (2: second step never streamed), 3 / 4 survivor queue
                 // drained 1 / 4 rows per lane at a time, 5 / 6 the same with
                 // the second step never streamed
};

// One probe step as the kernels see it. Codes of the step's dimension are
// packed 32/bits per word; the all-ones code is NULL.
struct StepDesc {
    const int32_t *fk;       // fact foreign-key column
    const uint32_t *words;   // packed codes (global copy)
    uint32_t lds_off;        // word offset of the LDS copy
    uint32_t in_lds;         // 1 when the kernel reads the LDS copy
    uint32_t dim_rows;       // keys >= dim_rows find no row
    uint32_t lw;             // log2(bits per code)
    uint32_t mask;           // NULL code = (1 << bits) - 1
    uint32_t factor;         // multiplier of this step's code
    uint32_t nwords;         // packed words of this step
    uint32_t pad_;
};

struct ScanArgs {
    StepDesc st[OLAP_MAX_DIMS];
    const int32_t *m1;
    const int32_t *m2;          // may be null
    uint64_t rows;
    uint64_t ntiles;
    unsigned long long *gsum;   // group_nums slots
    unsigned long long *gcnt;
    uint32_t *err;              // rows whose gid fell outside [0, G)
    uint32_t G;
    uint32_t gt_lds;            // group table in LDS
    uint32_t gt_off;            // word offset of the LDS group table (8-byte aligned)
    uint32_t list_off;          // word offset of the LDS tile list (compacting kernels)
    uint32_t count_only;        // no measure: only counts
    uint32_t eager1;            // row-wise: load step 1's keys with step 0's (most of its lines are live);
                                // selects the kernel variant, not read by the kernels
    // group table beyond LDS: an LDS cache of 2^hot_bits groups (0: none) at
    // word offset hot_off -- tags [S] (kHotEmpty or the group the slot
    // holds), counts [S], sums [S] (u64) -- that absorbs the matches of the
    // groups that claim a slot first; every other match takes global atomics
    uint32_t hot_off;
    uint32_t hot_bits;
    // column-wise state
    int32_t *gvec;              // per-row gid vector (static-vector column-wise)
    int32_t gvec_null;          // NULL value written into gvec
    int32_t step;               // probe step of a column-wise pass
    uint32_t last_step;         // column-wise static vector: final step writes gvec_null
    uint32_t *list_oid;         // per-tile segments (dynamic-vector column-wise)
    int32_t *list_gid;
    uint32_t *list_cnt;         // entries per tile
    StageArgs stage;            // staged aggregation passes
    // one-pass staging (AGG_CHUNK): the record pool and, at word offset
    // pc_lds, the per-partition chunk state (u32 current and spare chunk
    // ids, then u32 fill, chunk.nparts each)
    ChunkArgs chunk;
    uint32_t pc_lds;
    // row-wise survivor queues: word offset of the per-wave LDS queues
    uint32_t q_off;
    uint32_t launched;          // set by the row-wise launcher: OLAP_PIPE_* bits of the
                                // kernel it chose (not read by the kernels)
};

}  // namespace olap

// Context and query objects (opaque in the C ABI).
struct olap_ctx_s {
    int device = 0;
    int num_cus = 0;
    int64_t hbm_bytes = 0;
    char arch[64] = {0};
    std::string err;
    // autotuner cache: query class -> chosen launch shape and aggregation
    // layout of the row-wise scan
    struct Tuned {
        uint64_t key;
        int shape;
        int layout;
    };
    std::vector<Tuned> tuned;
    uint32_t *d_err = nullptr;     // error counter (4 words)
    void *d_scratch = nullptr;     // small reduction scratch
    size_t scratch_bytes = 0;
    // device address of this GPU's allocation-free scratch (olap_noalloc.hip)
    void *noalloc = nullptr;
    // join build: the row claiming each key, grown on demand
    int32_t *join_win = nullptr;
    size_t join_win_len = 0;
    // dimension programs: sorted IN values, grown on demand
    int32_t *in_vals = nullptr;
    size_t in_vals_len = 0;
    // vector_group beyond LDS: the staged form's scratch (grows only)
    olap::Stager *vg_stager = nullptr;
};

struct olap_query_s {
    olap_ctx ctx = nullptr;
    olap_plan plan{};
    olap_columns cols{};
    uint32_t *packed = nullptr;        // all steps' packed words
    bool owns_packed = false;          // packed was allocated for this query
    size_t packed_words = 0;
    olap::StepDesc steps[OLAP_MAX_DIMS]{};
    int64_t max_gid = 0;
    uint32_t *list_oid = nullptr;      // column-wise dynamic-vector lists
    int32_t *list_gid = nullptr;
    uint32_t *list_cnt = nullptr;
    size_t list_rows = 0;
    int32_t *gvec = nullptr;           // column-wise static vector scratch
    size_t gvec_rows = 0;
    bool joined_dv = false;
    int row_shape = -1;                // row-wise launch shape (-1: default / cached)
    // aggregation layout of a row-wise run whose group table exceeds LDS:
    // -1 automatic (from est_events), 0 global atomics, 1 staged
    int agg_layout = -1;
    double est_events = 0.0;           // expected matches (uniform keys assumed)
    uint32_t eager1 = 0;               // row-wise: stream step 1's keys with step 0's
    int32_t default_shape = 0;         // row-wise shape without a set or tuned one
    olap::Stager *stager = nullptr;
    uint32_t stage_grid = 0;
    // launch of the last run: threads per workgroup, rows per thread and the
    // aggregation (OLAP_AGG_*)
    int32_t ran_threads = 0, ran_rows = 0, ran_agg = -1;
    int32_t ran_pipe = 0;              // OLAP_PIPE_* bits of the last row-wise run
};

namespace olap {

// aggregation passes over caller-laid-out intermediates (olap_capi.cpp), used
// by the reference-signature layer: a per-row gid vector (NULL -1), or a flat
// (row, gid) list of n entries
olap_status agg_gid_vector(olap_ctx ctx, const int32_t *gvec, int64_t rows, const int32_t *m1,
                           const int32_t *m2, int64_t G, int64_t *sum, int64_t *count,
                           hipStream_t s);
olap_status agg_row_list(olap_ctx ctx, const uint32_t *oid, const int32_t *gid, int64_t n,
                         const int32_t *m1, const int32_t *m2, int64_t G, int64_t *sum,
                         int64_t *count, hipStream_t s);

// device address of the allocation-free scratch on `device`, which must be
// current; taken once per device (olap_noalloc.hip)
hipError_t noalloc_scratch_init(int device, void **ptr);

// launchers, olap_kernels.hip
hipError_t launch_zero_u64(unsigned long long *p, size_t n, hipStream_t s);
hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
hipError_t launch_zero_run(unsigned long long *sum, unsigned long long *count, size_t n, uint32_t *err,
                           hipStream_t s);
hipError_t launch_gen_uniform(int32_t *out, int64_t n, int64_t first, uint64_t key, int32_t lo,
                              int32_t hi, hipStream_t s);
hipError_t launch_gen_zipf(int32_t *out, int64_t n, int64_t first, uint64_t key,
                           const uint64_t *thr, int64_t d, hipStream_t s);
hipError_t launch_build_dimvec(const int32_t *filter, const int32_t *group, int64_t rows,
                               int32_t lo, int32_t hi, int32_t bitmap, int32_t group_mod,
                               int32_t code_bytes, void *out, hipStream_t s);
// stats[0] = max code (-1 when all NULL), stats[1] = negative non-NULL codes,
// stats[2] = non-NULL codes
hipError_t launch_dim_stats(const void *dimvec, int32_t code_bytes, int64_t rows,
                            int32_t *stats, hipStream_t s);
hipError_t launch_pack(const void *dimvec, int32_t code_bytes, int64_t rows, uint32_t lw,
                       uint32_t *words, uint32_t nwords, hipStream_t s);

// strategy kernels; args.gsum/gcnt must be zeroed by the caller
// one-pass staged row-wise scan (AGG_CHUNK): rows per tile (= records per
// pool chunk), the persistent grid, the launch
uint32_t rowwise_chunk_rows();
hipError_t rowwise_chunk_grid(int ndims, int num_cus, uint32_t dyn_lds, uint64_t rows, uint32_t *grid);
hipError_t run_rowwise_chunk(ScanArgs &a, int ndims, uint32_t dyn_lds, uint32_t grid, hipStream_t s);
hipError_t run_rowwise(ScanArgs &a, int ndims, int num_cus, uint32_t dyn_lds, bool vec,
                       int shape, hipStream_t s);
int num_row_shapes();
// row-wise shapes taken when none is set or tuned (survivor queue, 1024
// threads x 16 rows, 256-row drains; with / without step 1 streamed)
constexpr int kDefaultShapeStreamed = 11;
constexpr int kDefaultShapeUnstreamed = 15;
ScanShape row_shape(int i);
// launch shapes of the one-pass and the two-pass staged row-wise scans
ScanShape chunk_shape();
ScanShape staged_shape();
// staged row-wise scan (group table beyond LDS): grid shared by both passes
hipError_t rowwise_staged_grid(int ndims, int num_cus, uint32_t dyn_lds, uint64_t rows,
                               uint32_t *grid);
hipError_t run_rowwise_staged(ScanArgs &a, int ndims, uint32_t dyn_lds, int mode, uint32_t grid,
                              hipStream_t s);
hipError_t run_vwm_sv(ScanArgs &a, int ndims, int num_cus, uint32_t dyn_lds, bool vec,
                      hipStream_t s);
hipError_t run_vwm_dv(ScanArgs &a, int ndims, int num_cus, uint32_t dyn_lds, bool vec,
                      hipStream_t s);
hipError_t run_cwm_sv_join(ScanArgs &a, int num_cus, uint32_t dyn_lds, bool vec, hipStream_t s);
hipError_t run_cwm_sv_agg(ScanArgs &a, int num_cus, uint32_t dyn_lds, bool vec, hipStream_t s);
hipError_t run_cwm_dv_join(ScanArgs &a, int num_cus, uint32_t dyn_lds, bool vec, hipStream_t s);
hipError_t run_cwm_dv_agg(ScanArgs &a, int num_cus, uint32_t dyn_lds, hipStream_t s);

// vector_group: the group vector in LDS up to kVgLdsSlots groups, else
// global atomics (launch_vector_group); or, beyond LDS, the staged form: a
// COUNT and a SCATTER pass (mode AGG_COUNT / AGG_SCATTER) over one grid, then
// the Stager's reduce
constexpr int64_t kVgLdsSlots = 16384;
hipError_t launch_vector_group(int64_t rows, const int32_t *key1, const int32_t *m1,
                               const int32_t *m2, unsigned long long *gv, int64_t G,
                               unsigned long long *total, uint32_t *bad, int num_cus,
                               hipStream_t s);
hipError_t vector_group_staged_grid(int64_t rows, const int32_t *key1, const int32_t *m1, const int32_t *m2,
                                    uint32_t lds_bytes, int num_cus, uint32_t *grid);
hipError_t run_vector_group_staged(int mode, int64_t rows, const int32_t *key1, const int32_t *m1,
                                   const int32_t *m2, int64_t G, uint32_t *bad, const StageArgs &st,
                                   uint32_t lds_bytes, uint32_t grid, hipStream_t s);
// dimension programs (olap_predicate.hip); err counts rows whose code does
// not fit an int8 vector
hipError_t launch_dim_program(const DevProgram &p, int64_t rows, int32_t code_bytes, void *out,
                              uint32_t *err, hipStream_t s);
// validate a program and stage its IN values on stream s (olap_capi.cpp);
// the staged values stay valid until the next call on the context
olap_status prepare_program(olap_ctx ctx, const olap_dim_program *prog, int64_t rows,
                            DevProgram *out, hipStream_t s);
// TPC-H Q5's fused snowflake probe (olap_snowflake.hip)
hipError_t launch_tri_join(const int32_t *skey, int32_t sbias, const void *svec, int64_t srows,
                           int32_t sbytes, const int32_t *okey, int32_t obias, const void *ovec,
                           int64_t orows, int32_t obytes, const int32_t *price, const int32_t *disc,
                           int64_t rows, int32_t G, unsigned long long *sum, unsigned long long *cnt,
                           uint32_t *bad, uint32_t *sampled, int num_cus, hipStream_t s);  // sampled: 2 u32 of scratch
// dense-vector join sub-step (olap_join.hip); winner: vlen int32 of scratch
hipError_t launch_join_build(int64_t rows, const int32_t *key, const int32_t *payload,
                             int32_t *winner, int32_t *vec, int64_t vlen, uint32_t *bad,
                             int num_cus, hipStream_t s);
hipError_t launch_join_probe(int64_t rows, const int32_t *key, const int32_t *vec, int64_t vlen,
                             unsigned long long *total, uint32_t *bad, int num_cus, hipStream_t s);
hipError_t launch_truncate_u32(const unsigned long long *v, int64_t n, uint32_t *out, hipStream_t s);
hipError_t launch_sum_u64(const unsigned long long *v, int64_t n, unsigned long long *out,
                          int num_cus, hipStream_t s);
hipError_t launch_agg_expr(int64_t rows, const int32_t *p, const int32_t *q, const int32_t *t,
                           unsigned long long *out, int num_cus, hipStream_t s);

// Scratch of batched scans that outlives them: query masks, descriptors and
// the staged-aggregation buffers, grown on demand and never shrunk, so a
// caller running batch after batch (the server's workers) stops allocating
// and freeing device memory once warm. One batch at a time may use it.
struct BatchWorkspace;
BatchWorkspace *batch_workspace_create();
void batch_workspace_destroy(BatchWorkspace *ws);
// olap_batch_create on a borrowed workspace, with the mask builds and
// uploads ordered on stream s (the call synchronises s)
olap_status batch_create_in(olap_ctx ctx, const olap_star *star, const olap_batch_query *queries,
                            int32_t nq, BatchWorkspace *ws, hipStream_t s, olap_batch *out);

}  // namespace olap
