// olap_capi.cpp -- host side of the C ABI declared in include/olapcore.h.
//
// A query is created once per (plan, columns): the dimension vectors are
// validated on the device (max code, negative codes) and re-packed at the
// smallest power-of-two code width whose all-ones value is free for NULL.
// Each run then chooses the LDS residency of the packed vectors and of the
// group table for the requested strategy and launches persistent kernels.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "olap_internal.h"

using namespace olap;

namespace {

olap_status fail(olap_ctx ctx, olap_status st, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return st;
}

olap_status hip_fail(olap_ctx ctx, hipError_t e, const char *what)
{
    return fail(ctx, e == hipErrorOutOfMemory ? OLAP_E_NOMEM : OLAP_E_HIP, "%s: %s", what,
                hipGetErrorString(e));
}

#define HIPC(ctx, call)                                   \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

uint64_t host_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t stream_key(uint64_t seed, uint64_t stream)
{
    return host_mix64(seed ^ host_mix64(stream + 0x632be59bd9b4e019ULL));
}

void dev_free(void *p)
{
    if (p) (void)hipFree(p);
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

hipStream_t as_stream(void *s) { return (hipStream_t)s; }

uint32_t align_up(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }

// LDS residency of one run. Steps are placed in probe order (step 0 is read
// for every fact row), then the group table.
struct Layout {
    StepDesc st[OLAP_MAX_DIMS];
    uint32_t gt_lds = 0, gt_off = 0, list_off = 0, dyn_lds = 0;
};

Layout plan_layout(const olap_query_s *q, bool with_steps, bool with_table, bool with_list,
                   int only_step, uint32_t reserve = 0)
{
    Layout L;
    const int nd = q->plan.dimvec_nums;
    for (int s = 0; s < OLAP_MAX_DIMS; s++) {
        L.st[s] = q->steps[s];
        L.st[s].in_lds = 0;
        L.st[s].lds_off = 0;
    }
    uint32_t cap = kLdsBytes - (with_list ? kLdsListBytes : 0u) - 16u - reserve;
    uint32_t cur = 0;  // bytes
    const uint64_t G = (uint64_t)q->plan.group_nums;
    // With many matches (>= 1 in 256 rows) a group table that fits LDS on its
    // own takes precedence over the dimension vectors: each match would
    // otherwise cost a pair of global atomics, while a vector left in HBM
    // costs an L2 gather only for the rows that reach its step.
    uint32_t step_cap = cap;
    if (with_steps && with_table) {
        uint32_t steps_bytes = 0;
        for (int s = 0; s < nd; s++)
            if (only_step < 0 || s == only_step) steps_bytes += align_up(q->steps[s].nwords * 4u, 16u);
        const uint64_t tb = align_up((uint32_t)std::min<uint64_t>(G * 12u, 1u << 30), 16u);
        if (steps_bytes + tb > cap && tb <= cap && q->est_events * 256.0 >= (double)q->cols.rows)
            step_cap = cap - (uint32_t)tb;
    }
    if (with_steps) {
        for (int s = 0; s < nd; s++) {
            if (only_step >= 0 && s != only_step) continue;
            const uint32_t bytes = align_up(q->steps[s].nwords * 4u, 16u);
            if (cur + bytes <= step_cap) {
                L.st[s].in_lds = 1;
                L.st[s].lds_off = cur / 4u;
                cur += bytes;
            }
        }
    }
    if (with_table) {
        const uint64_t tb = G * 12u;
        const uint32_t off = align_up(cur, 16u);
        if (off + tb <= cap) {
            L.gt_lds = 1;
            L.gt_off = off / 4u;
            cur = off + (uint32_t)align_up((uint32_t)tb, 16u);
        }
    }
    if (with_list) {
        cur = align_up(cur, 16u);
        L.list_off = cur / 4u;
        cur += kLdsListBytes;
    }
    L.dyn_lds = std::max<uint32_t>(cur, 16u);
    return L;
}

void fill_args(ScanArgs &a, const olap_query_s *q, const Layout &L, int64_t *sum, int64_t *count)
{
    memset(&a, 0, sizeof(a));
    for (int s = 0; s < OLAP_MAX_DIMS; s++) a.st[s] = L.st[s];
    a.m1 = q->cols.m1;
    a.m2 = q->cols.m2;
    a.rows = (uint64_t)q->cols.rows;
    a.gsum = (unsigned long long *)sum;
    a.gcnt = (unsigned long long *)count;
    a.err = q->ctx->d_err;
    a.G = (uint32_t)q->plan.group_nums;
    a.gt_lds = L.gt_lds;
    a.gt_off = L.gt_off;
    a.list_off = L.list_off;
    a.count_only = q->cols.m1 == nullptr ? 1u : 0u;
    a.gvec_null = OLAP_GID_NULL;
    a.eager1 = q->eager1;
}

bool columns_vec_ok(const olap_query_s *q)
{
    for (int s = 0; s < q->plan.dimvec_nums; s++)
        if (!aligned16(q->steps[s].fk)) return false;
    if (q->cols.m1 && !aligned16(q->cols.m1)) return false;
    if (q->cols.m2 && !aligned16(q->cols.m2)) return false;
    return true;
}

olap_status ensure_gvec(olap_query_s *q)
{
    const size_t need = (size_t)q->cols.rows + 4;
    if (q->gvec && q->gvec_rows >= need) return OLAP_OK;
    if (q->gvec) dev_free(q->gvec);
    q->gvec = nullptr;
    HIPC(q->ctx, hipMalloc(&q->gvec, need * sizeof(int32_t)));
    q->gvec_rows = need;
    return OLAP_OK;
}

olap_status ensure_lists(olap_query_s *q)
{
    const size_t tiles = ((size_t)q->cols.rows + kTileV - 1) / kTileV;
    const size_t need = tiles * kTileV;
    if (q->list_oid && q->list_rows >= need) return OLAP_OK;
    if (q->list_oid) dev_free(q->list_oid);
    if (q->list_gid) dev_free(q->list_gid);
    if (q->list_cnt) dev_free(q->list_cnt);
    q->list_oid = nullptr;
    q->list_gid = nullptr;
    q->list_cnt = nullptr;
    HIPC(q->ctx, hipMalloc(&q->list_oid, std::max<size_t>(need, 1) * sizeof(uint32_t)));
    HIPC(q->ctx, hipMalloc(&q->list_gid, std::max<size_t>(need, 1) * sizeof(int32_t)));
    HIPC(q->ctx, hipMalloc(&q->list_cnt, std::max<size_t>(tiles, 1) * sizeof(uint32_t)));
    q->list_rows = need;
    return OLAP_OK;
}

olap_status zero_outputs(olap_query_s *q, int64_t *sum, int64_t *count, hipStream_t s)
{
    const size_t G = (size_t)q->plan.group_nums;
    HIPC(q->ctx, launch_zero_u64((unsigned long long *)sum, G, s));
    HIPC(q->ctx, launch_zero_u64((unsigned long long *)count, G, s));
    return OLAP_OK;
}

olap_status run_cwm_sv_joins(olap_query_s *q, int32_t *gvec, int32_t gvec_null, hipStream_t s)
{
    olap_ctx ctx = q->ctx;
    const int nd = q->plan.dimvec_nums;
    const bool vec = columns_vec_ok(q) && aligned16(gvec);
    for (int step = 0; step < nd; step++) {
        Layout L = plan_layout(q, true, false, false, step);
        if (L.st[step].in_lds) L.st[step].lds_off = 0;
        ScanArgs a;
        fill_args(a, q, L, nullptr, nullptr);
        a.gvec = gvec;
        a.step = step;
        a.last_step = step == nd - 1;
        a.gvec_null = gvec_null;
        HIPC(ctx, run_cwm_sv_join(a, ctx->num_cus, L.dyn_lds, vec, s));
    }
    return OLAP_OK;
}

olap_status run_cwm_sv_aggregate(olap_query_s *q, const int32_t *gvec, int64_t *sum,
                                 int64_t *count, hipStream_t s)
{
    Layout L = plan_layout(q, false, true, false, -1);
    ScanArgs a;
    fill_args(a, q, L, sum, count);
    a.gvec = const_cast<int32_t *>(gvec);
    const bool vec = columns_vec_ok(q) && aligned16(gvec);
    HIPC(q->ctx, run_cwm_sv_agg(a, q->ctx->num_cus, L.dyn_lds, vec, s));
    return OLAP_OK;
}

olap_status run_cwm_dv_joins(olap_query_s *q, hipStream_t s)
{
    olap_ctx ctx = q->ctx;
    olap_status st = ensure_lists(q);
    if (st != OLAP_OK) return st;
    const bool vec = columns_vec_ok(q);
    for (int step = 0; step < q->plan.dimvec_nums; step++) {
        Layout L = plan_layout(q, true, false, true, step);
        if (L.st[step].in_lds) L.st[step].lds_off = 0;
        ScanArgs a;
        fill_args(a, q, L, nullptr, nullptr);
        a.step = step;
        a.list_oid = q->list_oid;
        a.list_gid = q->list_gid;
        a.list_cnt = q->list_cnt;
        HIPC(ctx, run_cwm_dv_join(a, ctx->num_cus, L.dyn_lds, vec, s));
    }
    q->joined_dv = true;
    return OLAP_OK;
}

olap_status run_cwm_dv_aggregate(olap_query_s *q, int64_t *sum, int64_t *count, hipStream_t s)
{
    Layout L = plan_layout(q, false, true, false, -1);
    ScanArgs a;
    fill_args(a, q, L, sum, count);
    a.list_oid = q->list_oid;
    a.list_gid = q->list_gid;
    a.list_cnt = q->list_cnt;
    HIPC(q->ctx, run_cwm_dv_agg(a, q->ctx->num_cus, L.dyn_lds, s));
    return OLAP_OK;
}

// autotuner cache key: strategy, probe steps, LDS residency of the group
// table, log2 of the fact-row count
uint64_t query_class(const olap_query_s *q, olap_strategy strat)
{
    Layout L = plan_layout(q, true, true, strat == OLAP_VWM_DV, -1);
    uint64_t rows_log2 = 0;
    while ((1ull << rows_log2) < (uint64_t)std::max<int64_t>(q->cols.rows, 1)) rows_log2++;
    return ((uint64_t)strat << 40) | ((uint64_t)q->plan.dimvec_nums << 32) |
           ((uint64_t)L.gt_lds << 31) | rows_log2;
}

const olap_ctx_s::Tuned *tuned_for(const olap_query_s *q)
{
    const uint64_t key = query_class(q, OLAP_ROWWISE);
    for (const auto &t : q->ctx->tuned)
        if (t.key == key) return &t;
    return nullptr;
}

int effective_shape(const olap_query_s *q)
{
    if (q->row_shape >= 0) return q->row_shape;
    const olap_ctx_s::Tuned *t = tuned_for(q);
    return t ? t->shape : q->default_shape;
}

// aggregation layout of a row-wise run whose group table exceeds LDS: the
// query's own setting, else the autotuner's choice for the query's class,
// else -1 (automatic)
int effective_layout(const olap_query_s *q)
{
    if (q->agg_layout >= 0) return q->agg_layout;
    const olap_ctx_s::Tuned *t = tuned_for(q);
    return t ? t->layout : -1;
}

// Group table beyond LDS: staged aggregation when the expected matches make
// global atomics (~24 G/s) cost more than a second scan of the key columns.
bool auto_staged(const olap_query_s *q)
{
    return q->est_events > 2.0e6 && q->est_events > 0.03 * (double)q->cols.rows;
}

olap_status run_rowwise_stage(olap_query_s *q, int64_t *sum, int64_t *count, hipStream_t s)
{
    olap_ctx ctx = q->ctx;
    const int nd = q->plan.dimvec_nums;
    const uint32_t G = (uint32_t)q->plan.group_nums;
    const uint32_t nparts = (G + kPartSlots - 1) / kPartSlots;
    if (nparts > kMaxParts) return fail(ctx, OLAP_E_CAPACITY, "too many partitions");
    const uint32_t cnt_bytes = (nparts * 4u + 15u) & ~15u;
    const uint32_t stage_bytes = cnt_bytes + nparts * 8u;
    if (stage_bytes + 1024u > kLdsBytes)
        return fail(ctx, OLAP_E_CAPACITY, "staged aggregation: %u partitions exceed LDS", nparts);
    Layout L = plan_layout(q, true, false, false, -1, stage_bytes);
    const uint32_t cur_lds = align_up(L.dyn_lds, 16u) / 4u;
    const uint32_t start_lds = cur_lds + cnt_bytes / 4u;
    const uint32_t dyn = start_lds * 4u + nparts * 8u;
    uint32_t grid = 0;
    HIPC(ctx, rowwise_staged_grid(nd, ctx->num_cus, dyn, (uint64_t)q->cols.rows, &grid));
    if (!q->stager) q->stager = new Stager();
    Stager &S = *q->stager;
    if (S.nparts != nparts || S.nwg != grid) {
        const unsigned long long off0 = 0;
        HIPC(ctx, S.setup(&off0, &G, 1, grid, q->cols.m2 != nullptr, s));
    }
    ScanArgs a;
    fill_args(a, q, L, sum, count);
    a.stage = S.args(cur_lds, start_lds, (unsigned long long *)sum, (unsigned long long *)count);
    HIPC(ctx, run_rowwise_staged(a, nd, dyn, AGG_COUNT, grid, s));
    HIPC(ctx, S.prefix(s));
    unsigned long long total = 0;
    HIPC(ctx, S.total(&total, s));
    HIPC(ctx, S.reserve(total));
    a.stage = S.args(cur_lds, start_lds, (unsigned long long *)sum, (unsigned long long *)count);
    HIPC(ctx, run_rowwise_staged(a, nd, dyn, AGG_SCATTER, grid, s));
    HIPC(ctx, S.reduce((unsigned long long *)sum, (unsigned long long *)count, ctx->num_cus, s));
    const ScanShape sh = staged_shape();
    q->ran_threads = sh.bs;
    q->ran_rows = sh.unroll * kChunk;
    q->ran_agg = OLAP_AGG_TWOPASS;
    q->ran_pipe = 0;
    return OLAP_OK;
}

// one-pass staged aggregation of a group table beyond LDS: every match
// writes a record into its partition's pool chunk during the scan, and one
// reduce adds the records up (no count pass, no host read)
olap_status run_rowwise_onepass(olap_query_s *q, int64_t *sum, int64_t *count, hipStream_t s)
{
    olap_ctx ctx = q->ctx;
    const int nd = q->plan.dimvec_nums;
    const uint32_t G = (uint32_t)q->plan.group_nums;
    const uint32_t nparts = (G + kPartSlots - 1) / kPartSlots;
    if (nparts > kMaxParts) return fail(ctx, OLAP_E_CAPACITY, "too many partitions");
    const uint32_t pc_bytes = align_up(nparts * 12u, 16u);  // current, spare, fill per partition
    if (pc_bytes + 1024u > kLdsBytes)
        return fail(ctx, OLAP_E_CAPACITY, "one-pass staging: %u partitions exceed LDS", nparts);
    Layout L = plan_layout(q, true, false, false, -1, pc_bytes);
    const uint32_t pc_lds = align_up(L.dyn_lds, 16u) / 4u;
    const uint32_t dyn = pc_lds * 4u + pc_bytes;
    uint32_t grid = 0;
    HIPC(ctx, rowwise_chunk_grid(nd, ctx->num_cus, dyn, (uint64_t)q->cols.rows, &grid));
    if (!q->stager) q->stager = new Stager();
    Stager &S = *q->stager;
    if (S.nparts != nparts || S.nwg != grid || S.part_bits != kPartBits) {
        const unsigned long long off0 = 0;
        HIPC(ctx, S.setup(&off0, &G, 1, grid, q->cols.m2 != nullptr, s));
    }
    const uint32_t ch_bits = (uint32_t)__builtin_ctz(rowwise_chunk_rows());
    const uint32_t fmt = q->cols.m2 ? REC_WIDE : REC_NARROW;
    // every workgroup holds two chunks open per partition: refuse a reserve
    // above 1/32 of HBM (many partitions), which the two-pass form avoids
    const double open_bytes = 2.0 * nparts * grid * (double)rowwise_chunk_rows() * (fmt == REC_WIDE ? 16.0 : 6.0);
    if (open_bytes > (double)std::max<int64_t>(ctx->hbm_bytes, 1ll << 32) / 32.0)
        return fail(ctx, OLAP_E_CAPACITY, "one-pass staging: %u partitions x %u workgroups would hold %.1f GB of open chunks",
                    nparts, grid, open_bytes / 1e9);
    HIPC(ctx, S.chunk_reserve((unsigned long long)std::max(0.0, q->est_events), grid, ch_bits, fmt, 0));
    HIPC(ctx, S.chunk_clear(s));
    ScanArgs a;
    fill_args(a, q, L, sum, count);
    a.chunk = S.chunk_args();
    a.pc_lds = pc_lds;
    HIPC(ctx, run_rowwise_chunk(a, nd, dyn, grid, s));
    HIPC(ctx, S.chunk_reduce((unsigned long long *)sum, (unsigned long long *)count, ctx->num_cus, s));
    const ScanShape sh = chunk_shape();
    q->ran_threads = sh.bs;
    q->ran_rows = sh.unroll * kChunk;
    q->ran_agg = OLAP_AGG_ONEPASS;
    q->ran_pipe = 0;
    return OLAP_OK;
}

// argument checks of a query (plan and columns), before any device work
olap_status check_query_args(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols)
{
    const int nd = plan->dimvec_nums;
    if (nd < 1 || nd > OLAP_MAX_DIMS) return fail(ctx, OLAP_E_INVALID, "dimvec_nums must be 1..4");
    if (plan->group_nums < 1 || plan->group_nums > INT32_MAX)
        return fail(ctx, OLAP_E_INVALID, "group_nums must be in [1, 2^31)");
    if (cols->rows < 0 || cols->rows > (int64_t)UINT32_MAX - kTileRW)
        return fail(ctx, OLAP_E_CAPACITY, "rows must be in [0, 2^32 - %d)", kTileRW);
    bool seen[OLAP_MAX_DIMS] = {false, false, false, false};
    for (int s = 0; s < nd; s++) {
        const int t = plan->orders[s];
        if (t < 0 || t >= OLAP_MAX_DIMS || seen[t])
            return fail(ctx, OLAP_E_INVALID, "orders must be distinct table slots 0..3");
        seen[t] = true;
        if (plan->factor[s] < 0) return fail(ctx, OLAP_E_INVALID, "factor must be >= 0");
        if (cols->dim_rows[t] < 0 || cols->dim_rows[t] > INT32_MAX)
            return fail(ctx, OLAP_E_CAPACITY, "dim_rows of table %d out of range", t);
        if (cols->code_bytes[t] != 1 && cols->code_bytes[t] != 4)
            return fail(ctx, OLAP_E_INVALID, "code_bytes of table %d must be 1 or 4", t);
        if (cols->rows > 0 && !cols->fk[t])
            return fail(ctx, OLAP_E_INVALID, "fk column of table %d is null", t);
        if (cols->dim_rows[t] > 0 && !cols->dimvec[t])
            return fail(ctx, OLAP_E_INVALID, "dimension vector of table %d is null", t);
    }
    return OLAP_OK;
}

// Validate the dimension vectors of q (max code, negative codes, non-NULL
// codes per step), choose the probe order and pack the vectors. dstats:
// 12 bytes per step of device scratch. packed: where the packed words go
// (packed_cap words); null allocates them for q. Synchronises s.
olap_status query_prepare(olap_query_s *q, int32_t *dstats, uint32_t *packed, size_t packed_cap,
                          hipStream_t s0)
{
    olap_ctx ctx = q->ctx;
    const olap_plan *plan = &q->plan;
    const olap_columns *cols = &q->cols;
    const int nd = plan->dimvec_nums;
    // validation: max code, negative non-NULL codes and non-NULL codes per step
    int32_t hstats[3 * OLAP_MAX_DIMS];
    for (int s = 0; s < nd; s++) {
        hstats[3 * s] = -1;
        hstats[3 * s + 1] = 0;
        hstats[3 * s + 2] = 0;
    }
    hipError_t e = hipMemcpyAsync(dstats, hstats, 12 * nd, hipMemcpyHostToDevice, s0);
    for (int s = 0; s < nd && e == hipSuccess; s++) {
        const int t = plan->orders[s];
        e = launch_dim_stats(cols->dimvec[t], cols->code_bytes[t], cols->dim_rows[t],
                             dstats + 3 * s, s0);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hstats, dstats, 12 * nd, hipMemcpyDeviceToHost, s0);
    if (e == hipSuccess) e = hipStreamSynchronize(s0);
    if (e != hipSuccess) return hip_fail(ctx, e, "query_create: validation");
    int64_t max_gid = 0;
    uint32_t lw[OLAP_MAX_DIMS] = {0, 0, 0, 0};
    uint32_t nwords[OLAP_MAX_DIMS] = {0, 0, 0, 0};
    size_t total_words = 0;
    double pass = (double)cols->rows;
    for (int s = 0; s < nd; s++) {
        const int t = plan->orders[s];
        if (hstats[3 * s + 1] != 0)
            return fail(ctx, OLAP_E_INVALID, "table %d holds %d negative non-NULL codes", t,
                        hstats[3 * s + 1]);
        pass *= cols->dim_rows[t] > 0 ? (double)hstats[3 * s + 2] / (double)cols->dim_rows[t] : 0.0;
        const int64_t mx = hstats[3 * s];
        if (mx > 0) max_gid += mx * (int64_t)plan->factor[s];
        // smallest power-of-two width whose all-ones value exceeds every code
        uint32_t l = 0;
        while (l < 5 && ((1ull << (1u << l)) - 1ull) <= (uint64_t)(mx < 0 ? 0 : mx)) l++;
        lw[s] = l;
        const uint64_t per = 32u >> l;
        nwords[s] = (uint32_t)(((uint64_t)cols->dim_rows[t] + per - 1) / per);
        total_words += align_up(nwords[s], 4u);
    }
    if (max_gid >= plan->group_nums)
        return fail(ctx, OLAP_E_INVALID,
                    "largest reachable gid %lld is outside group_nums %lld (factor/codes mismatch)",
                    (long long)max_gid, (long long)plan->group_nums);
    q->max_gid = max_gid;
    q->est_events = pass;
    if (packed) {
        if (total_words > packed_cap)
            return fail(ctx, OLAP_E_CAPACITY,
                        "packed dimension vectors take %zu bytes, beyond the %zu-byte scratch",
                        total_words * 4, packed_cap * 4);
        q->packed = packed;
        q->owns_packed = false;
    } else {
        e = hipMalloc(&q->packed, std::max<size_t>(total_words, 4) * 4);
        if (e != hipSuccess) {
            q->packed = nullptr;
            return hip_fail(ctx, e, "query_create: packed vectors");
        }
        q->owns_packed = true;
    }
    q->packed_words = total_words;
    // Probe order. A row passes only if every step passes and its gid is the
    // sum of code * factor over the steps, so the order changes no result; it
    // decides how many later foreign-key lines the early exit still fetches
    // and how many rows reach each probe. Independent filters are best taken
    // by falling (1 - pass fraction) / cost: a step whose packed vector is
    // larger than LDS stays in HBM and costs an L2 gather per row that
    // reaches it (cost 4 against 1 for an LDS probe), so a selective step in
    // LDS can go before a slightly more selective one in HBM. With every
    // vector in LDS this is the most selective step first.
    int perm[OLAP_MAX_DIMS];
    double frac[OLAP_MAX_DIMS], rank[OLAP_MAX_DIMS];
    for (int s = 0; s < nd; s++) {
        const int t = plan->orders[s];
        perm[s] = s;
        frac[s] = cols->dim_rows[t] > 0 ? (double)hstats[3 * s + 2] / (double)cols->dim_rows[t] : 0.0;
        const bool hbm = (uint64_t)nwords[s] * 4u > (uint64_t)kLdsBytes;
        rank[s] = (1.0 - frac[s]) / (hbm ? 4.0 : 1.0);
    }
    std::stable_sort(perm, perm + nd, [&](int x, int y) {
        return rank[x] != rank[y] ? rank[x] > rank[y] : frac[x] < frac[y];
    });
    // step 1 is fetched with step 0 when most of its 128-byte lines (32 keys)
    // hold a row that passed step 0: a streamed column then costs less than
    // the dependent partial fetches
    q->eager1 = nd > 1 && 1.0 - std::pow(1.0 - frac[perm[0]], 32.0) > 0.5 ? 1u : 0u;
    // launch shape when none is set or tuned: the survivor-queue kernel, 1024
    // threads x 16 rows, 256-row drains (queues that do not fit take the
    // pipelined kernel). It streams step 1 only when nearly all of that
    // column's lines hold a survivor: its gathers fetch just the live lines,
    // which beats streaming well past the one-tile kernels' threshold (the
    // 600M-row 4-dim star, 73% live lines: 0.77 against 0.90 ms).
    q->default_shape = nd > 1 && 1.0 - std::pow(1.0 - frac[perm[0]], 32.0) > 0.95 ? kDefaultShapeStreamed
                                                                                 : kDefaultShapeUnstreamed;
    size_t off = 0;
    for (int s = 0; s < OLAP_MAX_DIMS; s++) memset(&q->steps[s], 0, sizeof(StepDesc));
    for (int i = 0; i < nd && e == hipSuccess; i++) {
        const int s = perm[i];
        const int t = plan->orders[s];
        StepDesc &st = q->steps[i];
        st.fk = cols->fk[t];
        st.words = q->packed + off;
        st.dim_rows = (uint32_t)cols->dim_rows[t];
        st.lw = lw[s];
        st.mask = lw[s] == 5 ? 0xffffffffu : ((1u << (1u << lw[s])) - 1u);
        st.factor = (uint32_t)plan->factor[s];
        st.nwords = nwords[s];
        e = launch_pack(cols->dimvec[t], cols->code_bytes[t], cols->dim_rows[t], lw[s],
                        q->packed + off, nwords[s], s0);
        off += align_up(nwords[s], 4u);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s0);
    if (e != hipSuccess) {
        if (q->owns_packed) dev_free(q->packed);
        q->packed = nullptr;
        return hip_fail(ctx, e, "query_create: pack");
    }
    return OLAP_OK;
}


struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

olap_status one_shot(olap_ctx ctx, olap_strategy strat, const olap_plan *plan,
                     const olap_columns *cols, int64_t *sum, int64_t *count, int64_t *total,
                     void *stream)
{
    olap_query q = nullptr;
    olap_status st = olap_query_create(ctx, plan, cols, &q);
    if (st != OLAP_OK) return st;
    st = olap_query_run(q, strat, sum, count, stream);
    if (st == OLAP_OK) {
        DeviceGuard g(ctx->device);
        hipStream_t s = as_stream(stream);
        unsigned long long *tot = (unsigned long long *)ctx->d_scratch;
        hipError_t e = launch_zero_u64(tot, 1, s);
        if (e == hipSuccess)
            e = launch_sum_u64((const unsigned long long *)count, plan->group_nums, tot,
                               ctx->num_cus, s);
        uint32_t herr = 0;
        unsigned long long htot = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&htot, tot, 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&herr, ctx->d_err, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(ctx, e, "one-shot finish");
        else if (herr) st = fail(ctx, OLAP_E_INVALID, "%u rows produced a gid outside group_nums", herr);
        else if (total) *total = (int64_t)htot;
    }
    olap_query_destroy(q);
    return st;
}

}  // namespace

namespace olap {

olap_status agg_gid_vector(olap_ctx ctx, const int32_t *gvec, int64_t rows, const int32_t *m1,
                           const int32_t *m2, int64_t G, int64_t *sum, int64_t *count,
                           hipStream_t s)
{
    olap_query_s q;
    q.ctx = ctx;
    q.plan.group_nums = G;
    q.cols.m1 = m1;
    q.cols.m2 = m2;
    q.cols.rows = rows;
    olap_status st = zero_outputs(&q, sum, count, s);
    if (st != OLAP_OK) return st;
    return run_cwm_sv_aggregate(&q, gvec, sum, count, s);
}

olap_status agg_row_list(olap_ctx ctx, const uint32_t *oid, const int32_t *gid, int64_t n,
                         const int32_t *m1, const int32_t *m2, int64_t G, int64_t *sum,
                         int64_t *count, hipStream_t s)
{
    // lay the list out as full kTileV-entry segments; the last one is short
    const int64_t tiles = (n + kTileV - 1) / kTileV;
    std::vector<uint32_t> cnt((size_t)std::max<int64_t>(tiles, 1));
    for (int64_t t = 0; t < tiles; t++)
        cnt[(size_t)t] = (uint32_t)std::min<int64_t>(kTileV, n - t * kTileV);
    uint32_t *dcnt = nullptr;
    HIPC(ctx, hipMalloc(&dcnt, cnt.size() * 4));
    hipError_t e = hipMemcpyAsync(dcnt, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, s);
    olap_query_s q;
    q.ctx = ctx;
    q.plan.group_nums = G;
    q.cols.m1 = m1;
    q.cols.m2 = m2;
    q.cols.rows = n;
    q.list_oid = const_cast<uint32_t *>(oid);
    q.list_gid = const_cast<int32_t *>(gid);
    q.list_cnt = dcnt;
    olap_status st = e == hipSuccess ? zero_outputs(&q, sum, count, s) : hip_fail(ctx, e, "agg_row_list");
    if (st == OLAP_OK) st = run_cwm_dv_aggregate(&q, sum, count, s);
    if (st == OLAP_OK) {
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(ctx, e, "agg_row_list sync");
    }
    q.list_oid = nullptr;  // caller-owned
    q.list_gid = nullptr;
    q.list_cnt = nullptr;
    dev_free(dcnt);
    return st;
}

// Validate a dimension program and stage its IN values (sorted, one
// device copy for all terms) on stream s.
olap_status prepare_program(olap_ctx ctx, const olap_dim_program *prog, int64_t rows,
                            DevProgram *out, hipStream_t s)
{
    if (!prog) return fail(ctx, OLAP_E_INVALID, "dimension program: null");
    if (prog->nterms < 1 || prog->nterms > OLAP_MAX_TERMS)
        return fail(ctx, OLAP_E_INVALID, "dimension program: 1..%d terms", OLAP_MAX_TERMS);
    memset(out, 0, sizeof(*out));
    out->nterms = prog->nterms;
    out->group = prog->group;
    out->group_mod = prog->group_mod;
    out->group_term = -1;
    out->group_col = prog->group_col;
    if (prog->group < OLAP_GROUP_NONE || prog->group > OLAP_GROUP_HOP)
        return fail(ctx, OLAP_E_INVALID, "dimension program: unknown group kind %d", prog->group);
    if (prog->group == OLAP_GROUP_COLUMN && ((rows && !prog->group_col) || prog->group_mod < 0))
        return fail(ctx, OLAP_E_INVALID, "dimension program: group column and group_mod >= 0 required");
    if (prog->group == OLAP_GROUP_HOP) {
        const int gt = prog->group_term;
        if (gt < 0 || gt >= prog->nterms || prog->term[gt].kind != OLAP_TERM_HOP || prog->term[gt].negate)
            return fail(ctx, OLAP_E_INVALID, "dimension program: group_term must name a HOP term that is not negated");
        out->group_term = gt;
    }
    std::vector<int32_t> vals;
    std::vector<size_t> voff(OLAP_MAX_TERMS, 0);
    for (int i = 0; i < prog->nterms; i++) {
        const olap_term &t = prog->term[i];
        DevTerm &d = out->term[i];
        d.kind = t.kind;
        d.combine = t.combine;
        d.negate = t.negate ? 1 : 0;
        d.lo = t.lo;
        d.hi = t.hi;
        d.key_bias = t.key_bias;
        d.col = t.col;
        if (t.combine != OLAP_AND && t.combine != OLAP_OR)
            return fail(ctx, OLAP_E_INVALID, "term %d: combine must be OLAP_AND or OLAP_OR", i);
        if (rows && !t.col) return fail(ctx, OLAP_E_INVALID, "term %d: null column", i);
        if (t.kind == OLAP_TERM_RANGE) {
            continue;
        } else if (t.kind == OLAP_TERM_IN) {
            if (t.nvalues < 1 || t.nvalues > OLAP_MAX_IN_VALUES || !t.values)
                return fail(ctx, OLAP_E_INVALID, "term %d: IN needs 1..%d host values", i, OLAP_MAX_IN_VALUES);
            std::vector<int32_t> v(t.values, t.values + t.nvalues);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            voff[(size_t)i] = vals.size();
            d.nvalues = (int32_t)v.size();
            vals.insert(vals.end(), v.begin(), v.end());
        } else if (t.kind == OLAP_TERM_HOP) {
            if ((t.parent_code_bytes != 1 && t.parent_code_bytes != 4) || t.parent_rows < 0 ||
                (t.parent_rows > 0 && !t.parent))
                return fail(ctx, OLAP_E_INVALID, "term %d: HOP needs a parent vector of 1- or 4-byte codes", i);
            d.parent = t.parent;
            d.parent_rows = t.parent_rows;
            d.parent_code_bytes = t.parent_code_bytes;
        } else {
            return fail(ctx, OLAP_E_INVALID, "term %d: unknown kind %d", i, t.kind);
        }
    }
    if (!vals.empty()) {
        if (vals.size() > ctx->in_vals_len) {
            HIPC(ctx, hipStreamSynchronize(s));  // earlier programs on s may still read the old copy
            dev_free(ctx->in_vals);
            ctx->in_vals = nullptr;
            ctx->in_vals_len = 0;
            HIPC(ctx, hipMalloc(&ctx->in_vals, vals.size() * 4));
            ctx->in_vals_len = vals.size();
        }
        HIPC(ctx, hipMemcpyAsync(ctx->in_vals, vals.data(), vals.size() * 4, hipMemcpyHostToDevice, s));
        for (int i = 0; i < prog->nterms; i++)
            if (prog->term[i].kind == OLAP_TERM_IN) out->term[i].values = ctx->in_vals + voff[(size_t)i];
        // the host copy must outlive the upload
        HIPC(ctx, hipStreamSynchronize(s));
    }
    return OLAP_OK;
}

}  // namespace olap

extern "C" {

olap_status olap_ctx_create(int device, olap_ctx *out)
{
    if (!out) return OLAP_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return OLAP_E_NODEVICE;
    if (device < 0 || device >= n) return OLAP_E_INVALID;
    olap_ctx ctx = new olap_ctx_s();
    ctx->device = device;
    DeviceGuard g(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete ctx;
        return OLAP_E_HIP;
    }
    ctx->num_cus = prop.multiProcessorCount;
    ctx->hbm_bytes = (int64_t)prop.totalGlobalMem;
    snprintf(ctx->arch, sizeof(ctx->arch), "%s", prop.gcnArchName);
    ctx->scratch_bytes = 256;
    if (hipMalloc(&ctx->d_err, 16) != hipSuccess ||
        hipMalloc(&ctx->d_scratch, ctx->scratch_bytes) != hipSuccess) {
        if (ctx->d_err) dev_free(ctx->d_err);
        delete ctx;
        return OLAP_E_NOMEM;
    }
    if (noalloc_scratch_init(device, &ctx->noalloc) != hipSuccess ||
        launch_zero_u32(ctx->d_err, 4, nullptr) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        dev_free(ctx->d_err);
        dev_free(ctx->d_scratch);
        delete ctx;
        return OLAP_E_HIP;
    }
    *out = ctx;
    return OLAP_OK;
}

olap_status olap_ctx_destroy(olap_ctx ctx)
{
    if (!ctx) return OLAP_E_INVALID;
    {
        DeviceGuard g(ctx->device);
        dev_free(ctx->d_err);
        dev_free(ctx->d_scratch);
        dev_free(ctx->join_win);
        dev_free(ctx->in_vals);
        delete ctx->vg_stager;
    }
    delete ctx;
    return OLAP_OK;
}

const char *olap_last_error(olap_ctx ctx) { return ctx ? ctx->err.c_str() : "no context"; }

olap_status olap_device_info(olap_ctx ctx, int32_t *num_cus, int64_t *hbm_bytes, char *arch_name,
                             size_t arch_len)
{
    if (!ctx) return OLAP_E_INVALID;
    if (num_cus) *num_cus = ctx->num_cus;
    if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
    if (arch_name && arch_len) snprintf(arch_name, arch_len, "%s", ctx->arch);
    return OLAP_OK;
}

olap_status olap_dev_alloc(olap_ctx ctx, size_t bytes, void **out)
{
    if (!ctx || !out) return OLAP_E_INVALID;
    DeviceGuard g(ctx->device);
    *out = nullptr;
    HIPC(ctx, hipMalloc(out, bytes ? bytes : 1));
    return OLAP_OK;
}

olap_status olap_dev_free(olap_ctx ctx, void *p)
{
    if (!ctx) return OLAP_E_INVALID;
    DeviceGuard g(ctx->device);
    if (p) HIPC(ctx, hipFree(p));
    return OLAP_OK;
}

olap_status olap_copy_h2d(olap_ctx ctx, void *dst, const void *src, size_t bytes, void *stream)
{
    if (!ctx || (bytes && (!dst || !src))) return OLAP_E_INVALID;
    DeviceGuard g(ctx->device);
    HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    HIPC(ctx, hipStreamSynchronize(as_stream(stream)));
    return OLAP_OK;
}

olap_status olap_copy_d2h(olap_ctx ctx, void *dst, const void *src, size_t bytes, void *stream)
{
    if (!ctx || (bytes && (!dst || !src))) return OLAP_E_INVALID;
    DeviceGuard g(ctx->device);
    HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    HIPC(ctx, hipStreamSynchronize(as_stream(stream)));
    return OLAP_OK;
}

olap_status olap_stream_sync(olap_ctx ctx, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    DeviceGuard g(ctx->device);
    HIPC(ctx, hipStreamSynchronize(as_stream(stream)));
    return OLAP_OK;
}

olap_status olap_gen_uniform_i32(olap_ctx ctx, int32_t *out, int64_t n, int64_t first_row,
                                 uint64_t seed, uint64_t stream_id, int32_t lo, int32_t hi,
                                 void *stream)
{
    if (!ctx || n < 0 || (n && !out) || first_row < 0) return fail(ctx, OLAP_E_INVALID, "gen_uniform: bad arguments");
    if ((int64_t)hi <= (int64_t)lo) return fail(ctx, OLAP_E_INVALID, "gen_uniform: empty range");
    DeviceGuard g(ctx->device);
    HIPC(ctx, launch_gen_uniform(out, n, first_row, stream_key(seed, stream_id), lo, hi,
                                 as_stream(stream)));
    return OLAP_OK;
}

olap_status olap_zipf_table(int64_t d, double theta, uint64_t *thr)
{
    if (d <= 0 || !thr || theta < 0.0) return OLAP_E_INVALID;
    double total = 0.0;
    for (int64_t k = 0; k < d; k++) total += 1.0 / pow((double)(k + 1), theta);
    double acc = 0.0;
    for (int64_t k = 0; k < d; k++) {
        acc += 1.0 / pow((double)(k + 1), theta);
        double v = floor(acc / total * 4294967296.0);
        if (v > 4294967296.0) v = 4294967296.0;
        thr[k] = (uint64_t)v;
    }
    thr[d - 1] = 4294967296ULL;
    return OLAP_OK;
}

olap_status olap_gen_zipf_i32(olap_ctx ctx, int32_t *out, int64_t n, int64_t first_row,
                              uint64_t seed, uint64_t stream_id, const uint64_t *thr_dev,
                              int64_t d, void *stream)
{
    if (!ctx || n < 0 || (n && (!out || !thr_dev)) || d <= 0 || d > INT32_MAX)
        return fail(ctx, OLAP_E_INVALID, "gen_zipf: bad arguments");
    DeviceGuard g(ctx->device);
    HIPC(ctx, launch_gen_zipf(out, n, first_row, stream_key(seed, stream_id), thr_dev, d,
                              as_stream(stream)));
    return OLAP_OK;
}

// sort_by_rate (test_multi_compute_operator.cpp:223-233): the comparison reads
// sele[] by position while only orders[] moves -- kept, since it fixes the
// reference's slot layout. Then main's group_nums / factor loop (:1977-1987).
olap_status olap_make_plan(const int32_t groups[OLAP_MAX_DIMS], const int32_t bitmap[OLAP_MAX_DIMS],
                           const double sele[OLAP_MAX_DIMS], olap_plan *out)
{
    if (!groups || !bitmap || !sele || !out) return OLAP_E_INVALID;
    olap_plan p;
    memset(&p, 0, sizeof(p));
    int32_t n = 0;
    for (int32_t t = 0; t < OLAP_MAX_DIMS; t++) {
        if (groups[t] < 0) return OLAP_E_INVALID;
        if (groups[t] != 0) p.orders[n++] = t;
    }
    if (n == 0) return OLAP_E_INVALID;
    p.dimvec_nums = n;
    for (int32_t i = 0; i < n; i++)
        for (int32_t j = i + 1; j < n; j++)
            if (sele[i] > sele[j]) std::swap(p.orders[i], p.orders[j]);
    int64_t G = 1;
    for (int32_t i = 0; i < OLAP_MAX_DIMS; i++) p.factor[i] = 1;
    for (int32_t i = 0; i < n; i++) {
        G *= groups[p.orders[i]];
        if (G > INT32_MAX) return OLAP_E_CAPACITY;
        int64_t f = 1;
        for (int32_t j = i + 1; j < n; j++) {
            if (bitmap[p.orders[i]] == 0) break;
            f *= groups[p.orders[j]];
        }
        p.factor[i] = (int32_t)f;
    }
    p.group_nums = G;
    *out = p;
    return OLAP_OK;
}

olap_status olap_build_dimvec(olap_ctx ctx, const int32_t *filter, const int32_t *group,
                              int64_t rows, int32_t lo, int32_t hi, int32_t bitmap,
                              int32_t group_mod, int32_t code_bytes, void *out, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (rows < 0 || (rows && (!filter || !out || (bitmap && !group))))
        return fail(ctx, OLAP_E_INVALID, "build_dimvec: bad arguments");
    if (code_bytes != 1 && code_bytes != 4)
        return fail(ctx, OLAP_E_INVALID, "build_dimvec: code_bytes must be 1 or 4");
    if (group_mod < 0) return fail(ctx, OLAP_E_INVALID, "build_dimvec: group_mod < 0");
    if (code_bytes == 1 && bitmap && (group_mod <= 0 || group_mod > OLAP_DIM_NULL_I8))
        return fail(ctx, OLAP_E_INVALID,
                    "build_dimvec: int8 codes need 0 < group_mod <= 127 to stay clear of NULL");
    DeviceGuard g(ctx->device);
    HIPC(ctx, launch_build_dimvec(filter, group, rows, lo, hi, bitmap, group_mod, code_bytes, out,
                                  as_stream(stream)));
    return OLAP_OK;
}

olap_status olap_build_dimvec_program(olap_ctx ctx, int64_t rows, const olap_dim_program *prog,
                                      int32_t code_bytes, void *out, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (rows < 0 || (rows && !out)) return fail(ctx, OLAP_E_INVALID, "build_dimvec_program: bad arguments");
    if (code_bytes != 1 && code_bytes != 4)
        return fail(ctx, OLAP_E_INVALID, "build_dimvec_program: code_bytes must be 1 or 4");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    DevProgram dp;
    olap_status st = prepare_program(ctx, prog, rows, &dp, s);
    if (st != OLAP_OK) return st;
    uint32_t *derr = (uint32_t *)ctx->d_scratch;
    HIPC(ctx, launch_zero_u32(derr, 1, s));
    HIPC(ctx, launch_dim_program(dp, rows, code_bytes, out, derr, s));
    uint32_t herr = 0;
    HIPC(ctx, hipMemcpyAsync(&herr, derr, 4, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipStreamSynchronize(s));
    if (herr)
        return fail(ctx, OLAP_E_INVALID, "build_dimvec_program: %u passing rows have codes outside 0..126 "
                                         "for an int8 vector", herr);
    return OLAP_OK;
}

olap_status olap_dict_encode(const char *const *strs, int64_t n, int32_t *codes, int64_t *ndistinct)
{
    if (n < 0 || (n && (!strs || !codes))) return OLAP_E_INVALID;
    for (int64_t i = 0; i < n; i++)
        if (!strs[i]) return OLAP_E_INVALID;
    std::vector<int64_t> idx((size_t)n);
    for (int64_t i = 0; i < n; i++) idx[(size_t)i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int64_t a, int64_t b) { return strcmp(strs[a], strs[b]) < 0; });
    int64_t rank = -1;
    for (size_t j = 0; j < idx.size(); j++) {
        if (j == 0 || strcmp(strs[idx[j - 1]], strs[idx[j]]) != 0) rank++;
        if (rank > INT32_MAX) return OLAP_E_CAPACITY;
        codes[idx[j]] = (int32_t)rank;
    }
    if (ndistinct) *ndistinct = rank + 1;
    return OLAP_OK;
}

int32_t olap_dict_code_of(const char *const *strs, const int32_t *codes, int64_t n, const char *value)
{
    if (!strs || !codes || !value) return -1;
    for (int64_t i = 0; i < n; i++)
        if (strs[i] && strcmp(strs[i], value) == 0) return codes[i];
    return -1;
}

olap_status olap_query_create(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                              olap_query *out)
{
    if (!ctx) return OLAP_E_INVALID;
    if (!plan || !cols || !out) return fail(ctx, OLAP_E_INVALID, "query_create: null argument");
    *out = nullptr;
    olap_status st = check_query_args(ctx, plan, cols);
    if (st != OLAP_OK) return st;
    DeviceGuard g(ctx->device);
    olap_query q = new olap_query_s();
    q->ctx = ctx;
    q->plan = *plan;
    q->cols = *cols;
    st = query_prepare(q, (int32_t *)ctx->d_scratch, nullptr, 0, nullptr);
    if (st != OLAP_OK) {
        delete q;
        return st;
    }
    *out = q;
    return OLAP_OK;
}

olap_status olap_query_run(olap_query q, olap_strategy strategy, int64_t *sum, int64_t *count,
                           void *stream)
{
    if (!q) return OLAP_E_INVALID;
    olap_ctx ctx = q->ctx;
    if (!count || (!sum && q->cols.m1)) return fail(ctx, OLAP_E_INVALID, "query_run: null output");
    if (q->cols.rows > 0 && !q->cols.m1 && sum)
        return fail(ctx, OLAP_E_INVALID, "query_run: sum requested without a measure");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    const int nd = q->plan.dimvec_nums;
    // a sum pointer is required by the kernels; counting-only runs reuse count
    int64_t *sum_p = sum ? sum : count;
    olap_status st = OLAP_OK;
    HIPC(ctx, launch_zero_run((unsigned long long *)sum, (unsigned long long *)count, (size_t)q->plan.group_nums,
                              ctx->d_err, s));
    const bool vec = columns_vec_ok(q);
    switch (strategy) {
    case OLAP_ROWWISE:
    case OLAP_VWM_SV:
    case OLAP_VWM_DV: {
        Layout L = plan_layout(q, true, true, strategy == OLAP_VWM_DV, -1);
        int layout = strategy == OLAP_ROWWISE ? effective_layout(q) : 0;
        const bool own = q->agg_layout >= 0;  // set on the query, not taken from the cache
        // the one-pass form keeps the dimension vectors in LDS, also where
        // the group table would have displaced them
        if (layout == 3 && vec) {
            st = run_rowwise_onepass(q, sum_p, count, s);
            if (st != OLAP_E_CAPACITY || own) return st;
            ctx->err.clear();  // a cached choice that does not fit this query
            layout = -1;
        }
        if (layout == 1 && !own && (L.gt_lds || !vec)) layout = -1;
        if (!L.gt_lds && vec && (layout == 1 || (layout < 0 && auto_staged(q)))) {
            // automatic choice: one pass where its open chunks fit, else two,
            // else (partition state beyond LDS) global atomics
            if (layout < 0) {
                st = run_rowwise_onepass(q, sum_p, count, s);
                if (st != OLAP_E_CAPACITY) return st;
                ctx->err.clear();
            }
            st = run_rowwise_stage(q, sum_p, count, s);
            if (st != OLAP_E_CAPACITY || own) return st;
            ctx->err.clear();
            layout = 0;
        }
        ScanArgs a;
        fill_args(a, q, L, sum_p, count);
        uint32_t dyn = L.dyn_lds;
        if (strategy == OLAP_ROWWISE && !L.gt_lds && layout == 2) {
            // global atomics behind an LDS cache of hot groups: 16 bytes a
            // slot, 256..4096 slots in the LDS the steps leave
            const uint32_t base = align_up(L.dyn_lds, 16u);
            const uint32_t avail = kLdsBytes - 16u - std::min(base, kLdsBytes - 16u);
            uint32_t bits = 12;
            while (bits > 8 && (16u << bits) > avail) bits--;
            if ((16u << bits) <= avail) {
                a.hot_off = base / 4u;
                a.hot_bits = bits;
                dyn = base + (16u << bits);
            }
        }
        hipError_t e;
        const ScanShape sh = strategy == OLAP_ROWWISE && vec ? row_shape(effective_shape(q))
                                                             : ScanShape{kBlock, kUnroll};
        if (strategy == OLAP_ROWWISE)
            e = run_rowwise(a, nd, ctx->num_cus, dyn, vec, effective_shape(q), s);
        else if (strategy == OLAP_VWM_SV) e = run_vwm_sv(a, nd, ctx->num_cus, L.dyn_lds, vec, s);
        else e = run_vwm_dv(a, nd, ctx->num_cus, L.dyn_lds, vec, s);
        if (e != hipSuccess) return hip_fail(ctx, e, "scan kernel");
        q->ran_threads = sh.bs;
        q->ran_rows = sh.unroll * kChunk;
        q->ran_agg = L.gt_lds ? OLAP_AGG_LDS : a.hot_bits ? OLAP_AGG_HOT : OLAP_AGG_ATOMICS;
        q->ran_pipe = strategy == OLAP_ROWWISE ? (int32_t)a.launched : 0;
        return OLAP_OK;
    }
    case OLAP_CWM_SV: {
        st = ensure_gvec(q);
        if (st != OLAP_OK) return st;
        st = run_cwm_sv_joins(q, q->gvec, OLAP_GID_NULL, s);
        if (st != OLAP_OK) return st;
        return run_cwm_sv_aggregate(q, q->gvec, sum_p, count, s);
    }
    case OLAP_CWM_DV: {
        st = run_cwm_dv_joins(q, s);
        if (st != OLAP_OK) return st;
        return run_cwm_dv_aggregate(q, sum_p, count, s);
    }
    }
    return fail(ctx, OLAP_E_INVALID, "unknown strategy %d", (int)strategy);
}

olap_status olap_query_destroy(olap_query q)
{
    if (!q) return OLAP_E_INVALID;
    DeviceGuard g(q->ctx->device);
    if (q->packed && q->owns_packed) dev_free(q->packed);
    if (q->list_oid) dev_free(q->list_oid);
    if (q->list_gid) dev_free(q->list_gid);
    if (q->list_cnt) dev_free(q->list_cnt);
    if (q->gvec) dev_free(q->gvec);
    delete q->stager;
    delete q;
    return OLAP_OK;
}

int32_t olap_num_shapes(olap_strategy strategy)
{
    return strategy == OLAP_ROWWISE ? (int32_t)num_row_shapes() : 1;
}

olap_status olap_shape_info(olap_strategy strategy, int32_t shape, int32_t *threads,
                            int32_t *rows_per_thread)
{
    if (shape < 0 || shape >= olap_num_shapes(strategy)) return OLAP_E_INVALID;
    ScanShape sh = strategy == OLAP_ROWWISE ? row_shape(shape) : ScanShape{kBlock, kUnroll};
    if (threads) *threads = sh.bs;
    if (rows_per_thread) *rows_per_thread = sh.unroll * kChunk;
    return OLAP_OK;
}

olap_status olap_query_set_shape(olap_query q, int32_t shape)
{
    if (!q) return OLAP_E_INVALID;
    if (shape < -1 || shape >= (int32_t)num_row_shapes())
        return fail(q->ctx, OLAP_E_INVALID, "shape %d out of range", shape);
    q->row_shape = shape;
    return OLAP_OK;
}

olap_status olap_query_set_aggregation(olap_query q, int32_t layout)
{
    if (!q) return OLAP_E_INVALID;
    if (layout < -1 || layout > 3) return fail(q->ctx, OLAP_E_INVALID, "layout must be -1, 0, 1, 2 or 3");
    q->agg_layout = layout;
    return OLAP_OK;
}

olap_status olap_query_autotune(olap_query q, olap_strategy strategy, int64_t *sum, int64_t *count,
                                void *stream, int32_t *shape_out, double *ms_out)
{
    if (!q) return OLAP_E_INVALID;
    olap_ctx ctx = q->ctx;
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    // candidates: every launch shape; a row-wise run whose group table exceeds
    // LDS also tries every shape behind the LDS cache of hot groups, and the
    // two-pass and one-pass staged aggregations (one fixed shape each,
    // skipped where their partition state or open chunks would not fit)
    std::vector<std::pair<int, int>> cands;  // (shape, aggregation layout)
    const bool table_in_lds = plan_layout(q, true, true, strategy == OLAP_VWM_DV, -1).gt_lds != 0;
    for (int i = 0; i < olap_num_shapes(strategy); i++)
        cands.push_back({i, strategy == OLAP_ROWWISE && !table_in_lds ? 0 : -1});
    if (strategy == OLAP_ROWWISE && !table_in_lds)  // the same behind an LDS cache of hot groups
        for (int i = 0; i < olap_num_shapes(strategy); i++) cands.push_back({i, 2});
    if (strategy == OLAP_ROWWISE && !table_in_lds && columns_vec_ok(q)) cands.push_back({0, 1});
    // one-pass staged, also where a group table in LDS pushed dimension
    // vectors out to HBM (the one-pass form keeps them in LDS)
    if (strategy == OLAP_ROWWISE && columns_vec_ok(q)) {
        const Layout with_table = plan_layout(q, true, true, false, -1);
        const Layout steps_only = plan_layout(q, true, false, false, -1);
        bool displaced = false;
        for (int st_i = 0; st_i < q->plan.dimvec_nums; st_i++)
            displaced |= steps_only.st[st_i].in_lds && !with_table.st[st_i].in_lds;
        if (!table_in_lds || displaced) cands.push_back({0, 3});
    }
    const int n = (int)cands.size();
    hipEvent_t e0, e1;
    HIPC(ctx, hipEventCreate(&e0));
    HIPC(ctx, hipEventCreate(&e1));
    int best = 0;
    double best_ms = 1e30;
    const int saved = q->row_shape;
    const int saved_layout = q->agg_layout;
    olap_status st = OLAP_OK;
    for (int i = 0; i < n && st == OLAP_OK; i++) {
        q->row_shape = cands[(size_t)i].first;
        q->agg_layout = cands[(size_t)i].second;
        st = olap_query_run(q, strategy, sum, count, stream);  // warm
        if (st == OLAP_E_CAPACITY && (cands[(size_t)i].second == 1 || cands[(size_t)i].second == 3)) {
            // a staged form whose partition state or open chunks do not fit
            st = OLAP_OK;
            ctx->err.clear();
            continue;
        }
        std::vector<double> t;
        for (int r = 0; r < 3 && st == OLAP_OK; r++) {
            hipError_t e = hipEventRecord(e0, s);
            if (e == hipSuccess) st = olap_query_run(q, strategy, sum, count, stream);
            if (e == hipSuccess && st == OLAP_OK) e = hipEventRecord(e1, s);
            if (e == hipSuccess && st == OLAP_OK) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess && st == OLAP_OK) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) st = hip_fail(ctx, e, "autotune timing");
            t.push_back(ms);
        }
        if (st != OLAP_OK) break;
        std::sort(t.begin(), t.end());
        if (t[1] < best_ms) {
            best_ms = t[1];
            best = i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != OLAP_OK) {
        q->row_shape = saved;
        q->agg_layout = saved_layout;
        return st;
    }
    q->row_shape = cands[(size_t)best].first;
    q->agg_layout = cands[(size_t)best].second;
    if (strategy == OLAP_ROWWISE) {
        const uint64_t key = query_class(q, strategy);
        bool found = false;
        for (auto &t : ctx->tuned)
            if (t.key == key) {
                t.shape = q->row_shape;
                t.layout = q->agg_layout;
                found = true;
            }
        if (!found) ctx->tuned.push_back({key, q->row_shape, q->agg_layout});
    }
    // shape index; num_shapes(strategy) stands for a staged aggregation
    if (shape_out)
        *shape_out = q->agg_layout == 1 || q->agg_layout == 3 ? olap_num_shapes(strategy) : q->row_shape;
    if (ms_out) *ms_out = best_ms;
    // leave the outputs holding a complete run of the chosen shape
    return olap_query_run(q, strategy, sum, count, stream);
}

olap_status olap_query_launch_info(olap_query q, int32_t *threads, int32_t *rows_per_thread,
                                   int32_t *aggregation)
{
    if (!q) return OLAP_E_INVALID;
    if (q->ran_agg < 0) return fail(q->ctx, OLAP_E_INVALID, "launch_info: the query has not run");
    if (threads) *threads = q->ran_threads;
    if (rows_per_thread) *rows_per_thread = q->ran_rows;
    if (aggregation) *aggregation = q->ran_agg;
    return OLAP_OK;
}

int32_t olap_query_pipeline(olap_query q)
{
    if (!q || q->ran_agg < 0) return -1;
    return q->ran_pipe;
}

int32_t olap_shape_pipeline(olap_strategy strategy, int32_t shape)
{
    if (strategy != OLAP_ROWWISE || shape < 0 || shape >= (int32_t)num_row_shapes()) return 0;
    return row_shape(shape).pipe;
}

int64_t olap_query_device_bytes(olap_query q)
{
    if (!q) return -1;
    return (int64_t)(q->packed_words * 4 + q->list_rows * 8 +
                     (q->list_rows / kTileV) * 4 + q->gvec_rows * 4);
}

olap_status olap_rowwise(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                         int64_t *sum, int64_t *count, int64_t *total_count, void *stream)
{
    return one_shot(ctx, OLAP_ROWWISE, plan, cols, sum, count, total_count, stream);
}
// olap_rowwise with every piece of device scratch in the GPU's
// allocation-free array: the statistics, the error word, the total and the
// packed dimension vectors. Group tables beyond LDS take global atomics
// (never a staged record buffer). Calls that overlap on one GPU share the
// array and race.
olap_status olap_rowwise_noalloc(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                                 int64_t *sum, int64_t *count, int64_t *total_count, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (!plan || !cols) return fail(ctx, OLAP_E_INVALID, "rowwise_noalloc: null argument");
    olap_status st = check_query_args(ctx, plan, cols);
    if (st != OLAP_OK) return st;
    if (!count || (!sum && cols->m1)) return fail(ctx, OLAP_E_INVALID, "rowwise_noalloc: null output");
    if (cols->rows > 0 && !cols->m1 && sum)
        return fail(ctx, OLAP_E_INVALID, "rowwise_noalloc: sum requested without a measure");
    if (!ctx->noalloc) return fail(ctx, OLAP_E_HIP, "rowwise_noalloc: no scratch on this device");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    unsigned char *base = (unsigned char *)ctx->noalloc;
    int32_t *dstats = (int32_t *)base;                                // 12 bytes per step
    unsigned long long *dtot = (unsigned long long *)(base + 64);    // total count
    uint32_t *derr = (uint32_t *)(base + 128);                        // error word
    uint32_t *packed = (uint32_t *)(base + kNoallocHeader);
    olap_query_s q;
    q.ctx = ctx;
    q.plan = *plan;
    q.cols = *cols;
    q.agg_layout = 0;
    st = query_prepare(&q, dstats, packed, (kNoallocScratchBytes - kNoallocHeader) / 4, s);
    if (st != OLAP_OK) return st;
    int64_t *sum_p = sum ? sum : count;
    if (sum) st = zero_outputs(&q, sum, count, s);
    else HIPC(ctx, launch_zero_u64((unsigned long long *)count, (size_t)plan->group_nums, s));
    if (st != OLAP_OK) return st;
    HIPC(ctx, launch_zero_u32(derr, 1, s));
    HIPC(ctx, launch_zero_u64(dtot, 1, s));
    Layout L = plan_layout(&q, true, true, false, -1);
    ScanArgs a;
    fill_args(a, &q, L, sum_p, count);
    a.err = derr;
    hipError_t e = run_rowwise(a, plan->dimvec_nums, ctx->num_cus, L.dyn_lds, columns_vec_ok(&q),
                               effective_shape(&q), s);
    if (e == hipSuccess)
        e = launch_sum_u64((const unsigned long long *)count, plan->group_nums, dtot, ctx->num_cus, s);
    uint32_t herr = 0;
    unsigned long long htot = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&htot, dtot, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, derr, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(ctx, e, "rowwise_noalloc");
    if (herr) return fail(ctx, OLAP_E_INVALID, "%u rows produced a gid outside group_nums", herr);
    if (total_count) *total_count = (int64_t)htot;
    return OLAP_OK;
}

olap_status olap_vwm_dv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream)
{
    return one_shot(ctx, OLAP_VWM_DV, plan, cols, sum, count, total_count, stream);
}
olap_status olap_vwm_sv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream)
{
    return one_shot(ctx, OLAP_VWM_SV, plan, cols, sum, count, total_count, stream);
}
olap_status olap_cwm_dv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream)
{
    return one_shot(ctx, OLAP_CWM_DV, plan, cols, sum, count, total_count, stream);
}
olap_status olap_cwm_sv(olap_ctx ctx, const olap_plan *plan, const olap_columns *cols,
                        int64_t *sum, int64_t *count, int64_t *total_count, void *stream)
{
    return one_shot(ctx, OLAP_CWM_SV, plan, cols, sum, count, total_count, stream);
}

olap_status olap_join_cwm_sv(olap_query q, int32_t *gid_vec, void *stream)
{
    if (!q) return OLAP_E_INVALID;
    if (q->cols.rows > 0 && !gid_vec) return fail(q->ctx, OLAP_E_INVALID, "join_cwm_sv: null gid_vec");
    DeviceGuard g(q->ctx->device);
    return run_cwm_sv_joins(q, gid_vec, OLAP_GID_NULL, as_stream(stream));
}

olap_status olap_agg_cwm_sv(olap_query q, const int32_t *gid_vec, int64_t *sum, int64_t *count,
                            void *stream)
{
    if (!q) return OLAP_E_INVALID;
    if (!sum || !count || (q->cols.rows > 0 && !gid_vec))
        return fail(q->ctx, OLAP_E_INVALID, "agg_cwm_sv: null argument");
    DeviceGuard g(q->ctx->device);
    hipStream_t s = as_stream(stream);
    olap_status st = zero_outputs(q, sum, count, s);
    if (st != OLAP_OK) return st;
    return run_cwm_sv_aggregate(q, gid_vec, sum, count, s);
}

olap_status olap_join_cwm_dv(olap_query q, void *stream)
{
    if (!q) return OLAP_E_INVALID;
    DeviceGuard g(q->ctx->device);
    return run_cwm_dv_joins(q, as_stream(stream));
}

olap_status olap_agg_cwm_dv(olap_query q, int64_t *sum, int64_t *count, void *stream)
{
    if (!q) return OLAP_E_INVALID;
    if (!sum || !count) return fail(q->ctx, OLAP_E_INVALID, "agg_cwm_dv: null output");
    if (!q->joined_dv) return fail(q->ctx, OLAP_E_INVALID, "agg_cwm_dv before join_cwm_dv");
    DeviceGuard g(q->ctx->device);
    hipStream_t s = as_stream(stream);
    olap_status st = zero_outputs(q, sum, count, s);
    if (st != OLAP_OK) return st;
    return run_cwm_dv_aggregate(q, sum, count, s);
}

int32_t olap_legacy_count(const int64_t *sum_host, int64_t group_nums)
{
    uint32_t acc = 0;
    for (int64_t g = 0; g < group_nums; g++) acc += (uint32_t)(uint64_t)sum_host[g];
    return (int32_t)acc;
}

olap_status olap_starjoin(olap_ctx ctx, olap_strategy strategy, int64_t rows,
                          const int8_t *const dimvec[4], const int64_t dim_rows[4],
                          const int32_t *const fk[4], int32_t *result, int64_t *count,
                          void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (!dimvec || !dim_rows || !fk || !count) return fail(ctx, OLAP_E_INVALID, "starjoin: null argument");
    olap_plan p;
    memset(&p, 0, sizeof(p));
    p.dimvec_nums = 4;
    olap_columns c;
    memset(&c, 0, sizeof(c));
    for (int t = 0; t < 4; t++) {
        p.orders[t] = t;
        c.dimvec[t] = dimvec[t];
        c.dim_rows[t] = dim_rows[t];
        c.code_bytes[t] = 1;
        c.fk[t] = fk[t];
    }
    c.rows = rows;
    c.m1 = nullptr;
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    olap_status st;
    if (result) {
        // per-row result: fixed 4-bit fields (test_starjoin_option.cpp:76,87,98,109)
        olap_plan pr = p;
        pr.factor[0] = 1 << 12;
        pr.factor[1] = 1 << 8;
        pr.factor[2] = 1 << 4;
        pr.factor[3] = 1;
        pr.group_nums = INT32_MAX;
        olap_query qr = nullptr;
        st = olap_query_create(ctx, &pr, &c, &qr);
        if (st != OLAP_OK) return st;
        st = run_cwm_sv_joins(qr, result, OLAP_STARJOIN_NULL, s);
        olap_query_destroy(qr);
        if (st != OLAP_OK) return st;
    }
    // the count: every fact row that passes lands in the single slot
    p.group_nums = 1;
    olap_query q = nullptr;
    st = olap_query_create(ctx, &p, &c, &q);
    if (st != OLAP_OK) return st;
    int64_t *dcnt = (int64_t *)ctx->d_scratch + 8;
    olap_strategy run_as = strategy == OLAP_CWM_SV && result ? OLAP_ROWWISE : strategy;
    st = olap_query_run(q, run_as, nullptr, dcnt, stream);
    if (st == OLAP_OK) {
        hipError_t e = hipMemcpyAsync(count, dcnt, 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(ctx, e, "starjoin count");
    }
    olap_query_destroy(q);
    return st;
}

olap_status olap_vector_group(olap_ctx ctx, int64_t rows, const int32_t *key1, const int32_t *m1,
                              const int32_t *m2, int64_t *group_vector, int64_t group_num,
                              int64_t *total, int64_t *bad_keys, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (rows < 0 || group_num < 1 || !group_vector || !total || (rows && (!key1 || !m1)))
        return fail(ctx, OLAP_E_INVALID, "vector_group: bad arguments");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    unsigned long long *dtot = (unsigned long long *)ctx->d_scratch;
    HIPC(ctx, launch_zero_u64((unsigned long long *)group_vector, (size_t)group_num, s));
    HIPC(ctx, launch_zero_u64(dtot, 1, s));
    HIPC(ctx, launch_zero_u32(ctx->d_err, 1, s));
    // beyond LDS: the staged form (count, scatter records, reduce per
    // partition in LDS) where its partition state fits LDS, else atomics
    const int64_t nparts = (group_num + kPartSlots - 1) / kPartSlots;
    const uint32_t cnt_bytes = align_up((uint32_t)std::min<int64_t>(nparts, kMaxParts) * 4u, 16u);
    const uint32_t stage_bytes = cnt_bytes + (uint32_t)std::min<int64_t>(nparts, kMaxParts) * 8u;
    if (rows > 0 && group_num > kVgLdsSlots && group_num <= UINT32_MAX && nparts <= kMaxParts &&
        stage_bytes + 1024u <= kLdsBytes) {
        uint32_t grid = 1;
        HIPC(ctx, vector_group_staged_grid(rows, key1, m1, m2, stage_bytes, ctx->num_cus, &grid));
        if (!ctx->vg_stager) ctx->vg_stager = new Stager();
        Stager &S = *ctx->vg_stager;
        const unsigned long long off0 = 0;
        const uint32_t G32 = (uint32_t)group_num;
        HIPC(ctx, S.setup(&off0, &G32, 1, grid, m2 != nullptr, s));
        unsigned long long *gv = (unsigned long long *)group_vector;
        HIPC(ctx, run_vector_group_staged(AGG_COUNT, rows, key1, m1, m2, group_num, ctx->d_err,
                                          S.args(0, cnt_bytes / 4u, gv, nullptr), stage_bytes, grid, s));
        HIPC(ctx, S.prefix(s));
        unsigned long long recs = 0;
        HIPC(ctx, S.total(&recs, s));
        HIPC(ctx, S.reserve(recs));
        HIPC(ctx, run_vector_group_staged(AGG_SCATTER, rows, key1, m1, m2, group_num, ctx->d_err,
                                          S.args(0, cnt_bytes / 4u, gv, nullptr), stage_bytes, grid, s));
        HIPC(ctx, S.reduce(gv, nullptr, ctx->num_cus, s));
    } else {
        HIPC(ctx, launch_vector_group(rows, key1, m1, m2, (unsigned long long *)group_vector, group_num,
                                      dtot, ctx->d_err, ctx->num_cus, s));
    }
    HIPC(ctx, launch_sum_u64((const unsigned long long *)group_vector, group_num, dtot,
                             ctx->num_cus, s));
    unsigned long long ht = 0;
    uint32_t hb = 0;
    HIPC(ctx, hipMemcpyAsync(&ht, dtot, 8, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipMemcpyAsync(&hb, ctx->d_err, 4, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipStreamSynchronize(s));
    *total = (int64_t)ht;
    if (bad_keys) *bad_keys = hb;
    return OLAP_OK;
}

olap_status olap_join_build(olap_ctx ctx, int64_t rows, const int32_t *key, const int32_t *payload,
                            int32_t *vector, int64_t vector_len, int64_t *bad_keys, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (rows < 0 || rows > INT32_MAX || vector_len < 0 || vector_len > INT32_MAX ||
        (rows && (!key || !payload)) || (vector_len && !vector))
        return fail(ctx, OLAP_E_INVALID, "join_build: rows and vector_len must be in [0, 2^31)");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    if (vector_len > 0 && (size_t)vector_len > ctx->join_win_len) {
        dev_free(ctx->join_win);
        ctx->join_win = nullptr;
        ctx->join_win_len = 0;
        HIPC(ctx, hipMalloc(&ctx->join_win, (size_t)vector_len * 4));
        ctx->join_win_len = (size_t)vector_len;
    }
    uint32_t *dbad = (uint32_t *)ctx->d_scratch;
    HIPC(ctx, launch_zero_u32(dbad, 1, s));
    HIPC(ctx, launch_join_build(rows, key, payload, ctx->join_win, vector, vector_len, dbad,
                                ctx->num_cus, s));
    uint32_t hbad = 0;
    HIPC(ctx, hipMemcpyAsync(&hbad, dbad, 4, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipStreamSynchronize(s));
    if (bad_keys) *bad_keys = hbad;
    return OLAP_OK;
}

olap_status olap_join_probe(olap_ctx ctx, int64_t rows, const int32_t *key, const int32_t *vector,
                            int64_t vector_len, int64_t *total, int64_t *bad_keys, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (rows < 0 || vector_len < 0 || vector_len > INT32_MAX || (rows && !key) ||
        (vector_len && !vector))
        return fail(ctx, OLAP_E_INVALID, "join_probe: bad arguments");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    unsigned long long *dtot = (unsigned long long *)ctx->d_scratch;
    uint32_t *dbad = (uint32_t *)((char *)ctx->d_scratch + 8);
    HIPC(ctx, launch_zero_u64(dtot, 1, s));
    HIPC(ctx, launch_zero_u32(dbad, 1, s));
    HIPC(ctx, launch_join_probe(rows, key, vector, vector_len, dtot, dbad, ctx->num_cus, s));
    unsigned long long htot = 0;
    uint32_t hbad = 0;
    HIPC(ctx, hipMemcpyAsync(&htot, dtot, 8, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipMemcpyAsync(&hbad, dbad, 4, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipStreamSynchronize(s));
    if (total) *total = (int64_t)htot;
    if (bad_keys) *bad_keys = hbad;
    return OLAP_OK;
}

olap_status olap_legacy_u32(olap_ctx ctx, const int64_t *sum, int64_t group_nums, uint32_t *u32_out,
                            void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (group_nums < 0 || (group_nums && (!sum || !u32_out)))
        return fail(ctx, OLAP_E_INVALID, "legacy_u32: bad arguments");
    DeviceGuard g(ctx->device);
    HIPC(ctx, launch_truncate_u32((const unsigned long long *)sum, group_nums, u32_out, as_stream(stream)));
    return OLAP_OK;
}

olap_status olap_tri_table_join(olap_ctx ctx, const olap_tri_join *j, int64_t *sum, int64_t *count,
                                void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (!j || !sum || !count) return fail(ctx, OLAP_E_INVALID, "tri_table_join: null argument");
    if (j->group_nums < 1 || j->group_nums > OLAP_TRI_MAX_GROUPS)
        return fail(ctx, OLAP_E_CAPACITY, "tri_table_join: group_nums must be in [1, %d]", OLAP_TRI_MAX_GROUPS);
    if (j->rows < 0) return fail(ctx, OLAP_E_INVALID, "tri_table_join: rows < 0");
    for (int d = 0; d < 2; d++) {
        if (j->code_bytes[d] != 1 && j->code_bytes[d] != 4)
            return fail(ctx, OLAP_E_INVALID, "tri_table_join: code_bytes must be 1 or 4");
        if (j->dim_rows[d] < 0 || (j->dim_rows[d] && !j->dimvec[d]) || (j->rows && !j->key[d]))
            return fail(ctx, OLAP_E_INVALID, "tri_table_join: keys and dimension vectors required");
    }
    if (j->rows && (!j->price || !j->discount))
        return fail(ctx, OLAP_E_INVALID, "tri_table_join: price and discount columns required");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    uint32_t *dbad = (uint32_t *)ctx->d_scratch;
    HIPC(ctx, launch_zero_u64((unsigned long long *)sum, (size_t)j->group_nums, s));
    HIPC(ctx, launch_zero_u64((unsigned long long *)count, (size_t)j->group_nums, s));
    HIPC(ctx, launch_zero_u32(dbad, 1, s));
    HIPC(ctx, launch_tri_join(j->key[0], j->key_bias[0], j->dimvec[0], j->dim_rows[0], j->code_bytes[0],
                              j->key[1], j->key_bias[1], j->dimvec[1], j->dim_rows[1], j->code_bytes[1],
                              j->price, j->discount, j->rows, j->group_nums,
                              (unsigned long long *)sum, (unsigned long long *)count, dbad,
                              (uint32_t *)((char *)ctx->d_scratch + 16), ctx->num_cus, s));
    uint32_t hbad = 0;
    HIPC(ctx, hipMemcpyAsync(&hbad, dbad, 4, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipStreamSynchronize(s));
    if (hbad)
        return fail(ctx, OLAP_E_INVALID, "tri_table_join: %u matching rows carry a code outside group_nums", hbad);
    return OLAP_OK;
}

olap_status olap_agg_expr(olap_ctx ctx, int64_t rows, const int32_t *p, const int32_t *q,
                          const int32_t *t, int64_t *result, void *stream)
{
    if (!ctx) return OLAP_E_INVALID;
    if (rows < 0 || !result || (rows && (!p || !q || !t)))
        return fail(ctx, OLAP_E_INVALID, "agg_expr: bad arguments");
    DeviceGuard g(ctx->device);
    hipStream_t s = as_stream(stream);
    unsigned long long *dtot = (unsigned long long *)ctx->d_scratch;
    HIPC(ctx, launch_zero_u64(dtot, 1, s));
    HIPC(ctx, launch_agg_expr(rows, p, q, t, dtot, ctx->num_cus, s));
    unsigned long long h = 0;
    HIPC(ctx, hipMemcpyAsync(&h, dtot, 8, hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipStreamSynchronize(s));
    *result = (int64_t)h;
    return OLAP_OK;
}

}  // extern "C"
