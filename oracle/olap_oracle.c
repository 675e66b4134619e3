/*
 * olap_oracle.c -- CPU restatement of the reference OLAPcore operator.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline). The product
 * library never links or calls this file. See olap_oracle.h for the shared
 * semantics. Citations are reference file:line.
 */
#define _GNU_SOURCE
#include "olap_oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* counter-based generator: SplitMix64 finaliser over (stream key + index)   */
/* ------------------------------------------------------------------------ */

uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t orc_stream_key(uint64_t seed, uint64_t stream)
{
    return orc_mix64(seed ^ orc_mix64(stream + 0x632be59bd9b4e019ULL));
}

static inline uint64_t orc_draw(uint64_t key, uint64_t i)
{
    return orc_mix64(key + (i + 1) * 0x9e3779b97f4a7c15ULL);
}

void orc_gen_uniform_i32(int32_t *out, int64_t n, int64_t first_row,
                         uint64_t seed, uint64_t stream, int32_t lo, int32_t hi)
{
    uint64_t key = orc_stream_key(seed, stream);
    uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo);
    for (int64_t i = 0; i < n; i++) {
        uint64_t x = orc_draw(key, (uint64_t)(first_row + i));
        out[i] = (int32_t)((int64_t)lo + (int64_t)(((x >> 32) * range) >> 32));
    }
}

/* Inverse-CDF table of a Zipf(theta) law over d ranks: thr[k] is
 * floor(2^32 * P(rank <= k)), thr[d-1] = 2^32. A draw u = top 32 bits maps to
 * the first k with thr[k] > u. */
void orc_zipf_table(uint64_t *thr, int64_t d, double theta)
{
    double total = 0.0;
    for (int64_t k = 0; k < d; k++)
        total += 1.0 / pow((double)(k + 1), theta);
    double acc = 0.0;
    for (int64_t k = 0; k < d; k++) {
        acc += 1.0 / pow((double)(k + 1), theta);
        double v = floor(acc / total * 4294967296.0);
        if (v > 4294967296.0) v = 4294967296.0;
        thr[k] = (uint64_t)v;
    }
    thr[d - 1] = 4294967296ULL;
}

void orc_gen_zipf_i32(int32_t *out, int64_t n, int64_t first_row,
                      uint64_t seed, uint64_t stream,
                      const uint64_t *thr, int64_t d)
{
    uint64_t key = orc_stream_key(seed, stream);
    for (int64_t i = 0; i < n; i++) {
        uint64_t u = orc_draw(key, (uint64_t)(first_row + i)) >> 32;
        int64_t a = 0, b = d - 1; /* first k in [a,b] with thr[k] > u */
        while (a < b) {
            int64_t m = (a + b) >> 1;
            if (thr[m] > u) b = m; else a = m + 1;
        }
        out[i] = (int32_t)a;
    }
}

/* ------------------------------------------------------------------------ */
/* reference-compatible generator: include/gendata_util.hpp                 */
/* ------------------------------------------------------------------------ */

/* size_of_table, include/gendata_util.hpp:152-166; tables 0..4 =
 * customer, supplier, part, date, lineorder; bases include/metadata.h:71-75 */
int64_t orc_size_of_table(int table, double sf)
{
    switch (table) {
    case 0: return (int64_t)(int)(30000.0 * sf);
    case 1: return (int64_t)(int)(2000.0 * sf);
    case 2: return (int64_t)(int)(200000.0 * (1.0 + log2(sf)));
    case 3: return 7 * 365;
    case 4: return (int64_t)(int)(6000000.0 * sf);
    }
    return 0;
}

void orc_srand(unsigned seed) { srand(seed); }

/* rand_x(double,double), gendata_util.hpp:194-197 */
static double orc_rand_unit(void)
{
    return 0.0 + (double)rand() / (double)(RAND_MAX / (1.0 - 0.0));
}

/* rand_x(int,int), gendata_util.hpp:184-187 */
static int orc_rand_range(int lo, int hi)
{
    return rand() % (hi - lo + 1) + lo;
}

/* gen_data, include/gendata_util.hpp:449-542: dimension order c, s, p, d;
 * per dimension row one rate draw, then a group draw only when it passes and
 * the dimension bears groups; FKs over [0, D] inclusive; M1 = M2 = 5. */
void orc_compat_gen_data(double c_sele, double s_sele, double p_sele, double d_sele,
                         double sf, const int bitmap[4], const int groups[4],
                         int8_t *dimvec[4], int32_t *fk[4], int32_t *m1, int32_t *m2)
{
    const double sele[4] = {c_sele, s_sele, p_sele, d_sele};
    int64_t dsz[4];
    for (int t = 0; t < 4; t++) dsz[t] = orc_size_of_table(t, sf);
    for (int t = 0; t < 4; t++) {
        for (int64_t i = 0; i < dsz[t]; i++) {
            if (bitmap[t])
                dimvec[t][i] = orc_rand_unit() <= sele[t]
                                   ? (int8_t)orc_rand_range(0, groups[t] - 1)
                                   : (int8_t)ORC_DIM_NULL_I8;
            else
                dimvec[t][i] = orc_rand_unit() <= sele[t] ? 0 : (int8_t)ORC_DIM_NULL_I8;
        }
    }
    int64_t n = orc_size_of_table(4, sf);
    for (int64_t i = 0; i < n; i++) {
        for (int t = 0; t < 4; t++)
            fk[t][i] = orc_rand_range(0, (int)dsz[t]);
        m1[i] = 5;
        m2[i] = 5;
    }
}

/* ------------------------------------------------------------------------ */
/* plan                                                                      */
/* ------------------------------------------------------------------------ */

/* sort_by_rate, test_multi_compute_operator.cpp:223-233. The comparison reads
 * sele[] by position while only orders[] is permuted; kept as is. */
void orc_sort_by_rate(const double *sele, int32_t *orders, int32_t n)
{
    for (int32_t i = 0; i < n; i++)
        for (int32_t j = i + 1; j < n; j++)
            if (sele[i] > sele[j]) {
                int32_t t = orders[i];
                orders[i] = orders[j];
                orders[j] = t;
            }
}

/* main's plan, test_multi_compute_operator.cpp:1964-1987 */
int orc_make_plan(const int32_t groups[4], const int32_t bitmap[4],
                  const double sele[4], orc_plan *out)
{
    memset(out, 0, sizeof(*out));
    int32_t n = 0;
    for (int32_t t = 0; t < 4; t++)
        if (groups[t] != 0) out->orders[n++] = t;
    out->dimvec_nums = n;
    orc_sort_by_rate(sele, out->orders, n);
    out->group_nums = 1;
    for (int32_t i = 0; i < 4; i++) out->factor[i] = 1;
    for (int32_t i = 0; i < n; i++) {
        out->group_nums *= groups[out->orders[i]];
        for (int32_t j = i + 1; j < n; j++) {
            if (bitmap[out->orders[i]] == 0) break;
            out->factor[i] *= groups[out->orders[j]];
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* filter -> dimension vector                                                */
/* ------------------------------------------------------------------------ */

/* dimvec[r] = pred(filter[r]) ? code : NULL, the dimension-vector form of
 * code/TPCH_Q5/Q5.cpp:921-944 (DimVec_S/o) and of gendata_util.hpp:473-527:
 * code is the group value when the dimension bears groups, else 0. */
void orc_build_dimvec_i8(const int32_t *filter, const int32_t *group, int64_t rows,
                         int32_t lo, int32_t hi, int32_t bitmap, int8_t *out)
{
    for (int64_t r = 0; r < rows; r++) {
        int pass = filter[r] >= lo && filter[r] < hi;
        out[r] = pass ? (int8_t)(bitmap ? group[r] : 0) : (int8_t)ORC_DIM_NULL_I8;
    }
}

void orc_build_dimvec_i32(const int32_t *filter, const int32_t *group, int64_t rows,
                          int32_t lo, int32_t hi, int32_t bitmap, int32_t group_mod,
                          int32_t *out)
{
    for (int64_t r = 0; r < rows; r++) {
        int pass = filter[r] >= lo && filter[r] < hi;
        int32_t code = 0;
        if (bitmap) code = group_mod > 0 ? (int32_t)((uint32_t)group[r] % (uint32_t)group_mod) : group[r];
        out[r] = pass ? code : ORC_DIM_NULL_I32;
    }
}

/* General dimension predicates. The reference chains select nodes over one
 * table, each writing a result vector that the next ANDs or ORs in
 * (code/TPCH_Q5_operator/q5_operator.cpp:1288-1550: less / less_equal /
 * more / more_equal / equal / non_equal value tests, string equality, and
 * equal_col, a lookup of the row's key in another table's result vector,
 * :1472-1506; Select_Option runs the chain, :1551-1596), and builds
 * snowflake dimension vectors through a second table
 * (code/TPCH_Q5/Q5.cpp:921-944: supplier -> nation -> region, and orders ->
 * customer -> nation with a date range). Restated: a term is a range
 * lo <= v < hi (each value test is one: v < c is [INT_MIN, c), v == c is
 * [c, c+1), v != c its negation), an IN list (string equality over
 * dictionary codes), or a hop parent[col + key_bias] != NULL; terms combine
 * left to right with AND / OR. Deliberate difference: the reference's OR
 * nodes keep a row when the previous node stored 0 (pre_bmp == 0,
 * :1329-1331), which is the previous node's pass flag only for string tests;
 * here OR is the boolean OR of the passes. The group code is 0, a group
 * column (mod group_mod), or the parent code of a hop term (equal_col_storecol2,
 * :1487-1506, and DimVec_S/o storing the nation key). */
static int orc_in_list(const int32_t *v, int32_t n, int32_t x)
{
    for (int32_t i = 0; i < n; i++)
        if (v[i] == x) return 1;
    return 0;
}

int64_t orc_build_dimvec_program(const orc_program *p, int64_t rows, int32_t code_bytes, void *out)
{
    int64_t bad = 0;
    for (int64_t r = 0; r < rows; r++) {
        int acc = 0, gfound = 0;
        int32_t gcode = 0;
        for (int i = 0; i < p->nterms; i++) {
            const orc_term *t = &p->term[i];
            int ok;
            if (t->kind == ORC_TERM_RANGE) {
                ok = t->col[r] >= t->lo && t->col[r] < t->hi;
            } else if (t->kind == ORC_TERM_IN) {
                ok = orc_in_list(t->values, t->nvalues, t->col[r]);
            } else {
                int64_t k = (int64_t)t->col[r] + t->key_bias;
                int32_t code = -1;
                if (k >= 0 && k < t->parent_rows) {
                    if (t->parent_code_bytes == 1) {
                        int8_t c = ((const int8_t *)t->parent)[k];
                        code = c == ORC_DIM_NULL_I8 ? -1 : c;
                    } else {
                        code = ((const int32_t *)t->parent)[k];
                    }
                }
                ok = code != -1;
                if (i == p->group_term && p->group == ORC_GROUP_HOP) {
                    gfound = ok;
                    gcode = code;
                }
            }
            if (t->negate) ok = !ok;
            acc = i == 0 ? ok : (t->combine == ORC_OR ? (acc || ok) : (acc && ok));
        }
        int32_t code = 0;
        if (p->group == ORC_GROUP_COLUMN) {
            int32_t g = p->group_col[r];
            code = p->group_mod > 0 ? (int32_t)((uint32_t)g % (uint32_t)p->group_mod) : g;
        } else if (p->group == ORC_GROUP_HOP) {
            acc = acc && gfound;
            code = gcode;
        }
        if (code_bytes == 1) {
            if (acc && (code < 0 || code >= ORC_DIM_NULL_I8)) {
                bad++;
                acc = 0;
            }
            ((int8_t *)out)[r] = acc ? (int8_t)code : (int8_t)ORC_DIM_NULL_I8;
        } else {
            ((int32_t *)out)[r] = acc ? code : ORC_DIM_NULL_I32;
        }
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* shared helpers                                                            */
/* ------------------------------------------------------------------------ */

#define ORC_NULL_CODE INT64_MIN

/* dimvec[fk] or ORC_NULL_CODE when NULL / no dimension row */
static inline int64_t orc_probe(const orc_columns *c, int t, int64_t row)
{
    int64_t k = c->fk[t][row];
    if (k < 0 || k >= c->dim_rows[t]) return ORC_NULL_CODE;
    if (c->code_bytes[t] == 1) {
        int8_t v = ((const int8_t *)c->dimvec[t])[k];
        return v == ORC_DIM_NULL_I8 ? ORC_NULL_CODE : (int64_t)v;
    }
    int32_t v = ((const int32_t *)c->dimvec[t])[k];
    return v == ORC_DIM_NULL_I32 ? ORC_NULL_CODE : (int64_t)v;
}

static inline int64_t orc_measure(const orc_columns *c, int64_t row)
{
    int64_t m = c->m1[row];
    if (c->m2) m += c->m2[row];
    return m;
}

typedef struct {
    const orc_plan *p;
    const orc_columns *c;
    int64_t start, num;
    int64_t *sum, *cnt;
    int64_t bad;
    int32_t *gid_vec; /* column-wise static vector, shared */
    int phase;        /* column-wise static vector: step index or -1 = agg */
} orc_task;

static inline void orc_add(orc_task *a, int64_t gid, int64_t row)
{
    if (gid < 0 || gid >= a->p->group_nums) { a->bad++; return; }
    a->sum[gid] += orc_measure(a->c, row);
    a->cnt[gid] += 1;
}

int orc_available_cpus(void)
{
    cpu_set_t s;
    CPU_ZERO(&s);
    if (sched_getaffinity(0, sizeof(s), &s) != 0) return 1;
    int n = CPU_COUNT(&s);
    return n > 0 ? n : 1;
}

/* Fork nthreads pinned workers over contiguous row ranges, as the reference
 * drivers do (e.g. test_multi_compute_operator.cpp:973-1037): thread j gets
 * [j*N/n, ...), the last one the remainder; private tables reduced serially. */
static int64_t orc_run(void *(*fn)(void *), const orc_plan *p, const orc_columns *c,
                       int64_t *sum, int64_t *cnt, int nthreads, int32_t *gid_vec, int phase)
{
    if (nthreads < 1) nthreads = 1;
    int64_t G = p->group_nums;
    orc_task *tasks = (orc_task *)calloc((size_t)nthreads, sizeof(orc_task));
    pthread_t *tid = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    int64_t per = c->rows / nthreads, left = c->rows;
    cpu_set_t allowed;
    CPU_ZERO(&allowed);
    sched_getaffinity(0, sizeof(allowed), &allowed);
    int cpus[CPU_SETSIZE], ncpu = 0;
    for (int i = 0; i < CPU_SETSIZE; i++)
        if (CPU_ISSET(i, &allowed)) cpus[ncpu++] = i;
    for (int j = 0; j < nthreads; j++) {
        orc_task *a = &tasks[j];
        a->p = p;
        a->c = c;
        a->start = per * j;
        a->num = (j == nthreads - 1) ? left : per;
        left -= per;
        a->gid_vec = gid_vec;
        a->phase = phase;
        if (sum) {
            a->sum = (int64_t *)calloc((size_t)(G > 0 ? G : 1), sizeof(int64_t));
            a->cnt = (int64_t *)calloc((size_t)(G > 0 ? G : 1), sizeof(int64_t));
        }
    }
    for (int j = 0; j < nthreads; j++) {
        pthread_attr_t attr;
        pthread_attr_init(&attr);
        if (nthreads > 1 && ncpu > 0) {
            cpu_set_t s;
            CPU_ZERO(&s);
            CPU_SET(cpus[j % ncpu], &s);
            pthread_attr_setaffinity_np(&attr, sizeof(s), &s);
        }
        if (nthreads == 1) {
            fn(&tasks[j]);
            tid[j] = 0;
        } else if (pthread_create(&tid[j], &attr, fn, &tasks[j]) != 0) {
            fn(&tasks[j]); /* degrade to inline run, never exit */
            tid[j] = 0;
        }
        pthread_attr_destroy(&attr);
    }
    int64_t bad = 0;
    for (int j = 0; j < nthreads; j++) {
        if (tid[j]) pthread_join(tid[j], NULL);
        bad += tasks[j].bad;
    }
    if (sum) {
        memset(sum, 0, sizeof(int64_t) * (size_t)G);
        memset(cnt, 0, sizeof(int64_t) * (size_t)G);
        for (int j = 0; j < nthreads; j++) {
            for (int64_t g = 0; g < G; g++) {
                sum[g] += tasks[j].sum[g];
                cnt[g] += tasks[j].cnt[g];
            }
            free(tasks[j].sum);
            free(tasks[j].cnt);
        }
    }
    free(tasks);
    free(tid);
    return bad;
}

/* ------------------------------------------------------------------------ */
/* row-wise: OLAPcore_row_thread, test_multi_compute_operator.cpp:329-367    */
/* ------------------------------------------------------------------------ */

static void *orc_row_thread(void *arg)
{
    orc_task *a = (orc_task *)arg;
    const orc_plan *p = a->p;
    for (int64_t i = 0; i < a->num; i++) {
        int64_t row = a->start + i;
        int64_t gid = 0;
        int pass = 1;
        for (int j = 0; j < p->dimvec_nums; j++) {
            int64_t v = orc_probe(a->c, p->orders[j], row);
            if (v == ORC_NULL_CODE) { pass = 0; break; } /* early exit */
            gid += v * p->factor[j];
        }
        /* the reference reads M1[i] with the thread-local i (:360); the
         * fact row is `row`, which is what is aggregated here */
        if (pass) orc_add(a, gid, row);
    }
    return NULL;
}

int64_t orc_rowwise(const orc_plan *p, const orc_columns *c,
                    int64_t *sum, int64_t *cnt, int nthreads)
{
    return orc_run(orc_row_thread, p, c, sum, cnt, nthreads, NULL, 0);
}

/* ------------------------------------------------------------------------ */
/* vector-wise, dynamic vector: OLAPcore_vwm_dv_thread :375-435              */
/* ------------------------------------------------------------------------ */

static void *orc_vwm_dv_thread(void *arg)
{
    orc_task *a = (orc_task *)arg;
    const orc_plan *p = a->p;
    int64_t oid[ORC_VEC_BLOCK];
    int64_t gidv[ORC_VEC_BLOCK];
    int64_t nblock = a->num / ORC_VEC_BLOCK;
    for (int64_t it = 0; it <= nblock; it++) {
        int64_t len = (it == nblock) ? a->num % ORC_VEC_BLOCK : ORC_VEC_BLOCK;
        int64_t n = 0;
        for (int s = 0; s < p->dimvec_nums; s++) {
            int t = p->orders[s];
            if (s == 0) {
                n = 0;
                for (int64_t j = 0; j < len; j++) {
                    int64_t row = a->start + it * ORC_VEC_BLOCK + j;
                    int64_t v = orc_probe(a->c, t, row);
                    if (v != ORC_NULL_CODE) {
                        oid[n] = row;
                        gidv[n] = v * p->factor[s];
                        n++;
                    }
                }
            } else {
                int64_t m = n;
                n = 0;
                for (int64_t j = 0; j < m; j++) {
                    int64_t row = oid[j];
                    int64_t v = orc_probe(a->c, t, row);
                    if (v != ORC_NULL_CODE) {
                        oid[n] = row;
                        gidv[n] = gidv[j] + v * p->factor[s];
                        n++;
                    }
                }
            }
        }
        for (int64_t j = 0; j < n; j++) orc_add(a, gidv[j], oid[j]);
    }
    return NULL;
}

int64_t orc_vwm_dv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads)
{
    return orc_run(orc_vwm_dv_thread, p, c, sum, cnt, nthreads, NULL, 0);
}

/* ------------------------------------------------------------------------ */
/* vector-wise, static vector: OLAPcore_vwm_sv_thread :509-568               */
/* ------------------------------------------------------------------------ */

static void *orc_vwm_sv_thread(void *arg)
{
    orc_task *a = (orc_task *)arg;
    const orc_plan *p = a->p;
    int64_t gidv[ORC_VEC_BLOCK];
    uint8_t live[ORC_VEC_BLOCK];
    int64_t nblock = a->num / ORC_VEC_BLOCK;
    for (int64_t it = 0; it <= nblock; it++) {
        int64_t len = (it == nblock) ? a->num % ORC_VEC_BLOCK : ORC_VEC_BLOCK;
        int64_t base = a->start + it * ORC_VEC_BLOCK;
        for (int s = 0; s < p->dimvec_nums; s++) {
            int t = p->orders[s];
            for (int64_t j = 0; j < len; j++) {
                int64_t v = orc_probe(a->c, t, base + j); /* every row probed */
                if (s == 0) {
                    live[j] = v != ORC_NULL_CODE;
                    gidv[j] = live[j] ? v * p->factor[s] : 0;
                } else if (live[j] && v != ORC_NULL_CODE) {
                    gidv[j] += v * p->factor[s];
                } else {
                    live[j] = 0;
                }
            }
        }
        for (int64_t j = 0; j < len; j++)
            if (live[j]) orc_add(a, gidv[j], base + j);
    }
    return NULL;
}

int64_t orc_vwm_sv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads)
{
    return orc_run(orc_vwm_sv_thread, p, c, sum, cnt, nthreads, NULL, 0);
}

/* ------------------------------------------------------------------------ */
/* column-wise, dynamic vector: join_cwm_dv_thread :240-283 then             */
/* agg_cwm_dv_thread :645-655. Each worker runs all join passes over its own */
/* partition (the reference forks once per pass; partitions are disjoint).   */
/* ------------------------------------------------------------------------ */

static void *orc_cwm_dv_thread(void *arg)
{
    orc_task *a = (orc_task *)arg;
    const orc_plan *p = a->p;
    int64_t cap = a->num > 0 ? a->num : 1;
    int64_t *oid = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
    int64_t *gidv = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
    int64_t n = 0;
    for (int s = 0; s < p->dimvec_nums; s++) {
        int t = p->orders[s];
        if (s == 0) {
            for (int64_t i = 0; i < a->num; i++) {
                int64_t row = a->start + i;
                int64_t v = orc_probe(a->c, t, row);
                if (v != ORC_NULL_CODE) {
                    oid[n] = row;
                    gidv[n] = v * p->factor[s];
                    n++;
                }
            }
        } else {
            int64_t m = n;
            n = 0;
            for (int64_t i = 0; i < m; i++) {
                int64_t row = oid[i];
                int64_t v = orc_probe(a->c, t, row);
                if (v != ORC_NULL_CODE) {
                    oid[n] = row;
                    gidv[n] = gidv[i] + v * p->factor[s];
                    n++;
                }
            }
        }
    }
    for (int64_t i = 0; i < n; i++) orc_add(a, gidv[i], oid[i]);
    free(oid);
    free(gidv);
    return NULL;
}

int64_t orc_cwm_dv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads)
{
    return orc_run(orc_cwm_dv_thread, p, c, sum, cnt, nthreads, NULL, 0);
}

/* ------------------------------------------------------------------------ */
/* column-wise, static vector: join_cwm_sv_thread :290-322 (one pass per     */
/* dimension over a shared gid vector) then agg_cwm_sv_thread :662-673       */
/* ------------------------------------------------------------------------ */

static void *orc_cwm_sv_thread(void *arg)
{
    orc_task *a = (orc_task *)arg;
    const orc_plan *p = a->p;
    int32_t *g = a->gid_vec;
    if (a->phase >= 0) {
        int s = a->phase, t = p->orders[s];
        for (int64_t i = 0; i < a->num; i++) {
            int64_t row = a->start + i;
            int64_t v = orc_probe(a->c, t, row);
            if (s == 0) {
                g[row] = v != ORC_NULL_CODE ? (int32_t)(v * p->factor[s]) : ORC_GID_NULL;
            } else if (g[row] != ORC_GID_NULL && v != ORC_NULL_CODE) {
                g[row] = (int32_t)(g[row] + v * p->factor[s]);
            } else {
                g[row] = ORC_GID_NULL;
            }
        }
    } else {
        for (int64_t i = 0; i < a->num; i++) {
            int64_t row = a->start + i;
            if (g[row] != ORC_GID_NULL) orc_add(a, g[row], row);
        }
    }
    return NULL;
}

void orc_join_cwm_sv(const orc_plan *p, const orc_columns *c, int32_t *gid_out)
{
    for (int s = 0; s < p->dimvec_nums; s++)
        orc_run(orc_cwm_sv_thread, p, c, NULL, NULL, 1, gid_out, s);
}

int64_t orc_cwm_sv(const orc_plan *p, const orc_columns *c,
                   int64_t *sum, int64_t *cnt, int nthreads)
{
    int32_t *g = (int32_t *)malloc(sizeof(int32_t) * (size_t)(c->rows > 0 ? c->rows : 1));
    for (int s = 0; s < p->dimvec_nums; s++)
        orc_run(orc_cwm_sv_thread, p, c, NULL, NULL, nthreads, g, s);
    int64_t bad = orc_run(orc_cwm_sv_thread, p, c, sum, cnt, nthreads, g, -1);
    free(g);
    return bad;
}

/* the int returned by every reference driver: the uint32 group vector summed
 * into an int (e.g. test_multi_compute_operator.cpp:1027-1036) */
int32_t orc_legacy_count(const int64_t *sum, int64_t group_nums)
{
    uint32_t acc = 0;
    for (int64_t g = 0; g < group_nums; g++) acc += (uint32_t)(uint64_t)sum[g];
    return (int32_t)acc;
}

/* ------------------------------------------------------------------------ */
/* sub-step operators                                                        */
/* ------------------------------------------------------------------------ */

/* starjoinalgo_rowwise / _cwm_sv, code/starjoin/test_starjoin_option.cpp:37-121:
 * count rows passing all four dimensions c, s, p, d; when result is given it
 * receives the 4-bit-field gid (c<<12 | s<<8 | p<<4 | d) or GROUP_NULL (32767,
 * include/metadata.h:70) per row. */
int64_t orc_starjoin_count(int64_t rows, const int8_t *const dimvec[4],
                           const int64_t dim_rows[4], const int32_t *const fk[4],
                           int32_t *result)
{
    int64_t count = 0;
    for (int64_t i = 0; i < rows; i++) {
        int32_t gid = 0;
        int pass = 1;
        for (int t = 0; t < 4; t++) {
            int64_t k = fk[t][i];
            if (k < 0 || k >= dim_rows[t] || dimvec[t][k] == ORC_DIM_NULL_I8) { pass = 0; break; }
            gid += ((int32_t)dimvec[t][k]) << (4 * (3 - t));
        }
        if (pass) count++;
        if (result) result[i] = pass ? gid : 32767;
    }
    return count;
}

/* vector_group, code/group/test_group_option.cpp:44-60: 1-based dense keys,
 * group_vector[key-1] += m1 + m2; returns the sum over the vector. Keys
 * outside [1, group_num] are skipped and counted in bad_keys. */
int64_t orc_vector_group(int64_t rows, const int32_t *key1, const int32_t *m1,
                         const int32_t *m2, int64_t *group_vector, int64_t group_num,
                         int64_t *bad_keys)
{
    int64_t bad = 0;
    memset(group_vector, 0, sizeof(int64_t) * (size_t)group_num);
    for (int64_t i = 0; i < rows; i++) {
        int64_t k = (int64_t)key1[i] - 1;
        if (k < 0 || k >= group_num) { bad++; continue; }
        group_vector[k] += (int64_t)m1[i] + (m2 ? (int64_t)m2[i] : 0);
    }
    int64_t total = 0;
    for (int64_t g = 0; g < group_num; g++) total += group_vector[g];
    if (bad_keys) *bad_keys = bad;
    return total;
}

/* vecjoin build, code/join/test_join_option.cpp:30-36: vector[key[i]] =
 * payload[i] over the build relation, in row order (a later row with the same
 * key overwrites an earlier one); entries no key names are left as they are.
 * Keys outside [0, vector_len) are skipped and counted in bad_keys. */
void orc_join_build(int64_t rows, const int32_t *key, const int32_t *payload, int32_t *vector,
                    int64_t vector_len, int64_t *bad_keys)
{
    int64_t bad = 0;
    for (int64_t i = 0; i < rows; i++) {
        if (key[i] < 0 || key[i] >= vector_len) { bad++; continue; }
        vector[key[i]] = payload[i];
    }
    if (bad_keys) *bad_keys = bad;
}

/* vecjoin probe, code/join/test_join_option.cpp:58-64 (and probe_prefetch,
 * :72-82, the same sum): count += vector[key[i]] over the probe relation,
 * here in int64. Keys outside [0, vector_len) add nothing and are counted in
 * bad_keys. */
int64_t orc_join_probe(int64_t rows, const int32_t *key, const int32_t *vector,
                       int64_t vector_len, int64_t *bad_keys)
{
    int64_t bad = 0, s = 0;
    for (int64_t i = 0; i < rows; i++) {
        if (key[i] < 0 || key[i] >= vector_len) { bad++; continue; }
        s += vector[key[i]];
    }
    if (bad_keys) *bad_keys = bad;
    return s;
}

/* Tri_table_join_vector_thread, code/TPCH_Q5/Q5.cpp:945-978: a lineitem row
 * counts when its supplier's code bitmap_S[l_suppkey - 1] is not -1 and
 * equals its order's code bitmap_o[l_orderkey_new] (the customer's nation),
 * and GrpVex[code] += l_extendedprice * (1 - l_discount). Restated in fixed
 * point: price in cents, discount in percent, sum[code] += price * (100 -
 * discount) in int64 (10^4 x the reference's double), count[code] += 1. Keys
 * (plus their bias) outside a vector find no row. Returns the matching rows
 * whose code is outside [0, G) (not aggregated). */
int64_t orc_tri_table_join(int64_t rows, const int32_t *skey, int32_t sbias, const int32_t *svec,
                           int64_t srows, const int32_t *okey, int32_t obias, const int32_t *ovec,
                           int64_t orows, const int32_t *price, const int32_t *disc, int32_t G,
                           int64_t *sum, int64_t *cnt)
{
    int64_t bad = 0;
    memset(sum, 0, sizeof(int64_t) * (size_t)G);
    memset(cnt, 0, sizeof(int64_t) * (size_t)G);
    for (int64_t i = 0; i < rows; i++) {
        int64_t ks = (int64_t)skey[i] + sbias, ko = (int64_t)okey[i] + obias;
        if (ks < 0 || ks >= srows || svec[ks] == -1) continue;
        if (ko < 0 || ko >= orows || ovec[ko] != svec[ks]) continue;
        int32_t g = svec[ks];
        if (g < 0 || g >= G) { bad++; continue; }
        sum[g] += (int64_t)price[i] * (int64_t)(100 - disc[i]);
        cnt[g] += 1;
    }
    return bad;
}

/* aggalgo_rowwise, code/aggregation/test_aggregation_option.cpp:28-37:
 * SUM(price*qty*tax + price*qty) with tax a percentage (gendata_util.hpp:757,
 * tax = rand_x(L_TAX_MIN, L_TAX_MAX)/100). Restated in fixed point with scale
 * 100 and int64 arithmetic: returns SUM(p*q*t_pct + 100*p*q), i.e. exactly
 * 100x the reference's double result when that is exact. */
int64_t orc_agg_expr(int64_t rows, const int32_t *p, const int32_t *q, const int32_t *t)
{
    int64_t s = 0;
    for (int64_t i = 0; i < rows; i++) {
        int64_t pq = (int64_t)p[i] * (int64_t)q[i];
        s += pq * (int64_t)t[i] + 100 * pq;
    }
    return s;
}
