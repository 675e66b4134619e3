/*
 * olap_oracle.h -- CPU restatement of the reference OLAPcore star-join
 * group-by operator, used ONLY as the checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only to check or to time the
 * CPU path beside the GPU. The product library (libolapcore.so) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates. Paths are
 * relative to the reference repository root.
 *
 * Semantics shared by all strategies (the bit-exact contract):
 *   - dimension vectors hold one code per dimension row: int8 with NULL 127
 *     (DIM_NULL, include/metadata.h:69) or int32 with NULL -1 (wide mode);
 *   - a fact row passes when every probed dimension returns a non-NULL code;
 *     a foreign key outside [0, dim_rows) finds no dimension row and fails
 *     (the reference reads past the vector there, gendata_util.hpp:530-534);
 *   - gid = sum_i code_i * factor[i] over the probe steps i;
 *   - sum[gid] += (int64)m1[row] + (int64)m2[row] (m2 may be NULL),
 *     count[gid] += 1. Accumulation is int64 (the reference's uint32 vector
 *     wraps, test_multi_compute_operator.cpp:989-990).
 *   - rows whose gid falls outside [0, group_nums) are not aggregated and
 *     are counted in the return value (the reference writes out of bounds).
 *
 * Parity pinning: see tests/test_oracle_*.py -- brute-force nested-loop join
 * on tiny stars, closed forms (bincount), cross-strategy equality, and the
 * compiled reference (oracle/_ref) on the reference's own generator.
 */
#ifndef OLAP_ORACLE_H
#define OLAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_DIMS 4
#define ORC_DIM_NULL_I8 127
#define ORC_DIM_NULL_I32 (-1)
#define ORC_GID_NULL (-1)
#define ORC_VEC_BLOCK 1024 /* size_v, include/metadata.h:68 */

typedef struct {
    int32_t dimvec_nums;        /* probe steps, 1..4 */
    int32_t orders[ORC_MAX_DIMS]; /* table slot probed at step i */
    int32_t factor[ORC_MAX_DIMS]; /* multiplier of step i's code */
    int64_t group_nums;         /* slots in the output */
} orc_plan;

typedef struct {
    const void *dimvec[ORC_MAX_DIMS]; /* per table slot */
    int64_t dim_rows[ORC_MAX_DIMS];
    int32_t code_bytes[ORC_MAX_DIMS]; /* 1 or 4 */
    const int32_t *fk[ORC_MAX_DIMS];  /* per table slot */
    const int32_t *m1;
    const int32_t *m2;                /* may be NULL */
    int64_t rows;
} orc_columns;

/* ---- counter-based generator (shared definition with the GPU generator) */
uint64_t orc_mix64(uint64_t z);
uint64_t orc_stream_key(uint64_t seed, uint64_t stream);
void orc_gen_uniform_i32(int32_t *out, int64_t n, int64_t first_row,
                         uint64_t seed, uint64_t stream, int32_t lo, int32_t hi);
void orc_zipf_table(uint64_t *thr, int64_t d, double theta);
void orc_gen_zipf_i32(int32_t *out, int64_t n, int64_t first_row,
                      uint64_t seed, uint64_t stream,
                      const uint64_t *thr, int64_t d);

/* ---- reference-compatible generator (glibc rand(), gendata_util.hpp) */
int64_t orc_size_of_table(int table, double sf);
void orc_srand(unsigned seed);
void orc_compat_gen_data(double c_sele, double s_sele, double p_sele, double d_sele,
                         double sf, const int bitmap[4], const int groups[4],
                         int8_t *dimvec[4], int32_t *fk[4], int32_t *m1, int32_t *m2);

/* ---- plan (test_multi_compute_operator.cpp:223-233, :1964-1987) */
void orc_sort_by_rate(const double *sele, int32_t *orders, int32_t n);
int orc_make_plan(const int32_t groups[4], const int32_t bitmap[4],
                  const double sele[4], orc_plan *out);

/* ---- filter -> dimension vector (north-star step 1) */
void orc_build_dimvec_i8(const int32_t *filter, const int32_t *group, int64_t rows,
                         int32_t lo, int32_t hi, int32_t bitmap, int8_t *out);
/* general dimension predicates: terms combined left to right */
enum { ORC_TERM_RANGE = 0, ORC_TERM_IN = 1, ORC_TERM_HOP = 2 };
enum { ORC_AND = 0, ORC_OR = 1 };
enum { ORC_GROUP_NONE = 0, ORC_GROUP_COLUMN = 1, ORC_GROUP_HOP = 2 };
typedef struct {
    int32_t kind, combine, negate, lo, hi, nvalues, parent_code_bytes, key_bias;
    const int32_t *col;       /* host */
    const int32_t *values;    /* IN, any order */
    const void *parent;       /* HOP */
    int64_t parent_rows;
} orc_term;
typedef struct {
    int32_t nterms, group, group_mod, group_term;
    const int32_t *group_col;
    orc_term term[8];
} orc_program;
/* returns the rows whose code does not fit an int8 vector (written NULL) */
int64_t orc_build_dimvec_program(const orc_program *p, int64_t rows, int32_t code_bytes, void *out);
void orc_build_dimvec_i32(const int32_t *filter, const int32_t *group, int64_t rows,
                          int32_t lo, int32_t hi, int32_t bitmap, int32_t group_mod,
                          int32_t *out);

/* ---- OLAPcore strategies; return rows with out-of-range gid */
int64_t orc_rowwise(const orc_plan *p, const orc_columns *c,
                    int64_t *sum, int64_t *cnt, int nthreads);
int64_t orc_vwm_dv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads);
int64_t orc_vwm_sv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads);
int64_t orc_cwm_dv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads);
int64_t orc_cwm_sv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads);
/* join_cwm_sv alone: per-row gid (or ORC_GID_NULL) */
void orc_join_cwm_sv(const orc_plan *p, const orc_columns *c, int32_t *gid_out);

/* the reference's returned total: int sum of the uint32 group vector */
int32_t orc_legacy_count(const int64_t *sum, int64_t group_nums);

/* ---- sub-step operators */
int64_t orc_starjoin_count(int64_t rows, const int8_t *const dimvec[4],
                           const int64_t dim_rows[4], const int32_t *const fk[4],
                           int32_t *result);
int64_t orc_vector_group(int64_t rows, const int32_t *key1, const int32_t *m1,
                         const int32_t *m2, int64_t *group_vector, int64_t group_num,
                         int64_t *bad_keys);
int64_t orc_agg_expr(int64_t rows, const int32_t *p, const int32_t *q, const int32_t *t);
int64_t orc_tri_table_join(int64_t rows, const int32_t *skey, int32_t sbias, const int32_t *svec,
                           int64_t srows, const int32_t *okey, int32_t obias, const int32_t *ovec,
                           int64_t orows, const int32_t *price, const int32_t *disc, int32_t G,
                           int64_t *sum, int64_t *cnt);
void orc_join_build(int64_t rows, const int32_t *key, const int32_t *payload, int32_t *vector,
                    int64_t vector_len, int64_t *bad_keys);
int64_t orc_join_probe(int64_t rows, const int32_t *key, const int32_t *vector,
                       int64_t vector_len, int64_t *bad_keys);

/* threads actually usable by this process (sched_getaffinity) */
int orc_available_cpus(void);

#ifdef __cplusplus
}
#endif
#endif
