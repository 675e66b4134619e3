// olapcore_compat.cpp -- the reference's OLAPcore operator signatures
// (include/olapcore_compat.hpp) on top of the C ABI. Host arrays in, host
// results out; every call uploads the columns it reads.
#include <hip/hip_runtime.h>

#include <dirent.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "olap_internal.h"
#include "olapcore_compat.hpp"

// The reference's NUMA argument structs (include/metadata.h:272-279): one
// per node, the node's replica of every dimension vector and its shard of
// every foreign-key column, by table slot.
struct Dimvec_array_numa {
    int8_t *dimvec[4];
};
struct Fk_array_numa {
    int32_t *fk[4];
};

namespace {

constexpr int16_t kGroupNull16 = 32767;  // GROUP_NULL, include/metadata.h:70

std::mutex g_mu;
olap_ctx g_ctx = nullptr;
int g_device = 0;
int64_t g_dim_rows[4] = {-1, -1, -1, -1};
std::string g_err;
std::vector<int64_t> g_sum, g_cnt;

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p && g_ctx) olap_dev_free(g_ctx, p);
    }
};

bool ensure_ctx()
{
    if (g_ctx) return true;
    olap_status st = olap_ctx_create(g_device, &g_ctx);
    if (st != OLAP_OK) {
        g_err = "olap_ctx_create failed with status " + std::to_string((int)st);
        g_ctx = nullptr;
        return false;
    }
    return true;
}

bool upload(DevBuf &b, const void *host, size_t bytes)
{
    if (olap_dev_alloc(g_ctx, bytes ? bytes : 4, &b.p) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return false;
    }
    if (bytes && olap_copy_h2d(g_ctx, b.p, host, bytes, nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return false;
    }
    return true;
}

// Fact rows as the caller holds them: one host array per column, or (NUMA
// drivers) one array per node shard, the shards in row order.
struct FactShards {
    int n = 1;                       // shards
    std::vector<int64_t> rows;       // rows per shard
    std::vector<int32_t *const *> fk;  // per shard: fk[table]
    std::vector<const int32_t *> m1, m2;
};

// upload one fact column, shard by shard, into one device array
bool upload_column(DevBuf &b, const FactShards &F, int64_t total,
                   const int32_t *(*pick)(const FactShards &, int, int), int table)
{
    if (olap_dev_alloc(g_ctx, total ? 4 * (size_t)total : 4, &b.p) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return false;
    }
    int64_t off = 0;
    for (int i = 0; i < F.n; i++) {
        const int32_t *src = pick(F, i, table);
        if (F.rows[(size_t)i] && olap_copy_h2d(g_ctx, (int32_t *)b.p + off, src,
                                               4 * (size_t)F.rows[(size_t)i], nullptr) != OLAP_OK) {
            g_err = olap_last_error(g_ctx);
            return false;
        }
        off += F.rows[(size_t)i];
    }
    return true;
}

const int32_t *pick_fk(const FactShards &F, int i, int t) { return F.fk[(size_t)i][t]; }
const int32_t *pick_m1(const FactShards &F, int i, int) { return F.m1[(size_t)i]; }
const int32_t *pick_m2(const FactShards &F, int i, int) { return F.m2[(size_t)i]; }

FactShards one_shard(int32_t **fk_array, int rows, const int32_t *M1, const int32_t *M2)
{
    FactShards F;
    F.rows = {rows};
    F.fk = {fk_array};
    F.m1 = {M1};
    F.m2 = {M2};
    return F;
}

int32_t legacy(const std::vector<int64_t> &sum)
{
    return olap_legacy_count(sum.data(), (int64_t)sum.size());
}

// Uploads the tables the plan probes and builds an olap_query over them.
struct Staged {
    DevBuf dv[4], fk[4], m1, m2;
    olap_plan plan{};
    olap_columns cols{};
    olap_query q = nullptr;
    ~Staged()
    {
        if (q) olap_query_destroy(q);
    }
};

bool stage(Staged &S, int8_t *const *dimvec_array, const FactShards &F, const int *orders, int ndims,
           const int *factor, int rows, int group_nums)
{
    const bool with_m1 = F.m1[0] != nullptr, with_m2 = F.m2[0] != nullptr;
    if (!ensure_ctx()) return false;
    if (ndims < 1 || ndims > 4 || rows < 0 || group_nums < 1) {
        g_err = "bad dimvec_nums / size_lineorder / group_nums";
        return false;
    }
    S.plan.dimvec_nums = ndims;
    S.plan.group_nums = group_nums;
    for (int i = 0; i < ndims; i++) {
        const int t = orders[i];
        if (t < 0 || t > 3) {
            g_err = "orders[] must name table slots 0..3";
            return false;
        }
        if (g_dim_rows[t] < 0) {
            g_err = "olapcore_compat_set_dim_rows was not called for table " + std::to_string(t);
            return false;
        }
        S.plan.orders[i] = t;
        S.plan.factor[i] = factor[i];
        if (!upload(S.dv[t], dimvec_array[t], (size_t)g_dim_rows[t]) ||
            !upload_column(S.fk[t], F, rows, pick_fk, t))
            return false;
        S.cols.dimvec[t] = S.dv[t].p;
        S.cols.dim_rows[t] = g_dim_rows[t];
        S.cols.code_bytes[t] = 1;
        S.cols.fk[t] = (const int32_t *)S.fk[t].p;
    }
    for (int t = 0; t < 4; t++)
        if (!S.cols.code_bytes[t]) S.cols.code_bytes[t] = 1;
    if (with_m1 && !upload_column(S.m1, F, rows, pick_m1, 0)) return false;
    if (with_m2 && !upload_column(S.m2, F, rows, pick_m2, 0)) return false;
    S.cols.m1 = (const int32_t *)S.m1.p;
    S.cols.m2 = (const int32_t *)S.m2.p;
    S.cols.rows = rows;
    if (olap_query_create(g_ctx, &S.plan, &S.cols, &S.q) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        S.q = nullptr;
        return false;
    }
    return true;
}

bool fetch_results(int64_t G, const DevBuf &sum, const DevBuf &cnt)
{
    g_sum.assign((size_t)G, 0);
    g_cnt.assign((size_t)G, 0);
    if (olap_copy_d2h(g_ctx, g_sum.data(), sum.p, 8 * (size_t)G, nullptr) != OLAP_OK ||
        olap_copy_d2h(g_ctx, g_cnt.data(), cnt.p, 8 * (size_t)G, nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return false;
    }
    return true;
}

int run_fused(olap_strategy strat, int8_t *const *dimvec_array, const FactShards &F, int *orders,
              int dimvec_nums, int *factor, int rows, int group_nums)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Staged S;
    if (!stage(S, dimvec_array, F, orders, dimvec_nums, factor, rows, group_nums))
        return -1;
    DevBuf sum, cnt;
    if (olap_dev_alloc(g_ctx, 8 * (size_t)group_nums, &sum.p) != OLAP_OK ||
        olap_dev_alloc(g_ctx, 8 * (size_t)group_nums, &cnt.p) != OLAP_OK ||
        olap_query_run(S.q, strat, (int64_t *)sum.p, (int64_t *)cnt.p, nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return -1;
    }
    if (!fetch_results(group_nums, sum, cnt)) return -1;
    return legacy(g_sum);
}

int g_numa_nodes = 0;  // 0: count the host's nodes

// numa_max_node() + 1 as libnuma reports it: the highest node number under
// /sys/devices/system/node, plus one
int host_numa_nodes()
{
    int hi = -1;
    if (DIR *d = opendir("/sys/devices/system/node")) {
        while (dirent *e = readdir(d)) {
            int k = -1;
            char tail = 0;
            if (sscanf(e->d_name, "node%d%c", &k, &tail) == 1 && k > hi) hi = k;
        }
        closedir(d);
    }
    return hi >= 0 ? hi + 1 : 1;
}

// The NUMA drivers' fact layout (test_multi_compute_operator.cpp:2017-2100):
// node i holds rows [i * (N / n), ...) of every fact column, the last node
// the remainder, and a replica of every dimension vector.
int run_numa(olap_strategy strat, Dimvec_array_numa *dv, Fk_array_numa *fk, int *orders,
             int dimvec_nums, int *factor, int rows, int group_nums, int32_t **M1_p,
             int32_t **M2_p)
{
    int n;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        n = g_numa_nodes > 0 ? g_numa_nodes : host_numa_nodes();
        if (!dv || !fk || !M1_p) {
            g_err = "NUMA driver: null dimension, foreign-key or measure array";
            return -1;
        }
    }
    FactShards F;
    F.n = n;
    for (int i = 0; i < n; i++) {
        F.rows.push_back(i == n - 1 ? rows - (int64_t)(n - 1) * (rows / n) : rows / n);
        F.fk.push_back(fk[i].fk);
        F.m1.push_back(M1_p[i]);
        F.m2.push_back(M2_p ? M2_p[i] : nullptr);
    }
    return run_fused(strat, dv[0].dimvec, F, orders, dimvec_nums, factor, rows, group_nums);
}

// thread partition of the reference drivers: thread j owns rows
// [j * (N / n), ...), the last thread the remainder
int partition_of(int64_t row, int64_t rows, int nthreads)
{
    const int64_t per = rows / nthreads;
    if (per == 0) return nthreads - 1;
    const int64_t j = row / per;
    return (int)std::min<int64_t>(j, nthreads - 1);
}

}  // namespace

void olapcore_compat_set_dim_rows(const int64_t rows[4])
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (int t = 0; t < 4; t++) g_dim_rows[t] = rows[t];
}

void olapcore_compat_set_numa_nodes(int nodes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_numa_nodes = nodes > 0 ? nodes : 0;
}

void olapcore_compat_set_device(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ctx && device != g_device) {
        olap_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
    g_device = device;
}

const char *olapcore_compat_last_error() { return g_err.c_str(); }
const int64_t *olapcore_compat_last_sums() { return g_sum.data(); }
const int64_t *olapcore_compat_last_counts() { return g_cnt.data(); }

int OLAPcore_rowwise(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                     int *factor, int &size_lineorder, int &group_nums, int32_t *M1, int32_t *M2,
                     int nthreads)
{
    (void)nthreads;
    return run_fused(OLAP_ROWWISE, dimvec_array, one_shard(fk_array, size_lineorder, M1, M2), orders,
                     dimvec_nums, factor, size_lineorder, group_nums);
}

int OLAPcore_vwm_dv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                    int *factor, int &size_lineorder, int &group_nums, int32_t *M1, int32_t *M2,
                    int nthreads)
{
    (void)nthreads;
    return run_fused(OLAP_VWM_DV, dimvec_array, one_shard(fk_array, size_lineorder, M1, M2), orders,
                     dimvec_nums, factor, size_lineorder, group_nums);
}

int OLAPcore_vwm_sv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                    int *factor, int &size_lineorder, int &group_nums, int32_t *M1, int32_t *M2,
                    int nthreads)
{
    (void)nthreads;
    return run_fused(OLAP_VWM_SV, dimvec_array, one_shard(fk_array, size_lineorder, M1, M2), orders,
                     dimvec_nums, factor, size_lineorder, group_nums);
}

int OLAPcore_vwm_dv_numa(Dimvec_array_numa *dimvec_array_numa, Fk_array_numa *fk_array_numa,
                         int *orders, int &dimvec_nums, int *factor, int &size_lineorder,
                         int &group_nums, int32_t **M1_p, int32_t **M2_p, int nthreads)
{
    (void)nthreads;
    return run_numa(OLAP_VWM_DV, dimvec_array_numa, fk_array_numa, orders, dimvec_nums, factor,
                    size_lineorder, group_nums, M1_p, M2_p);
}

int OLAPcore_vwm_sv_numa(Dimvec_array_numa *dimvec_array_numa, Fk_array_numa *fk_array_numa,
                         int *orders, int &dimvec_nums, int *factor, int &size_lineorder,
                         int &group_nums, int32_t **M1_p, int32_t **M2_p, int nthreads)
{
    (void)nthreads;
    return run_numa(OLAP_VWM_SV, dimvec_array_numa, fk_array_numa, orders, dimvec_nums, factor,
                    size_lineorder, group_nums, M1_p, M2_p);
}

void join_cwm_dv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                 int64_t **OID, int16_t **groupID, int *factor, int *index, int &size_lineorder,
                 int nthreads)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (nthreads < 1) nthreads = 1;
    for (int j = 0; j < nthreads; j++) index[j] = 0;
    Staged S;
    // group_nums only bounds the gid check here; the lists carry raw gids
    if (!stage(S, dimvec_array, one_shard(fk_array, size_lineorder, nullptr, nullptr), orders,
               dimvec_nums, factor, size_lineorder, INT32_MAX))
        return;
    if (olap_join_cwm_dv(S.q, nullptr) != OLAP_OK || olap_stream_sync(g_ctx, nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return;
    }
    olap_query_s *q = S.q;
    const size_t tiles = ((size_t)size_lineorder + olap::kTileV - 1) / olap::kTileV;
    std::vector<uint32_t> cnt(tiles), oid(q->list_rows);
    std::vector<int32_t> gid(q->list_rows);
    if (tiles && (hipMemcpy(cnt.data(), q->list_cnt, tiles * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(oid.data(), q->list_oid, q->list_rows * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(gid.data(), q->list_gid, q->list_rows * 4, hipMemcpyDeviceToHost) != hipSuccess)) {
        g_err = "join_cwm_dv: list download failed";
        return;
    }
    std::vector<std::vector<std::pair<int64_t, int32_t>>> part((size_t)nthreads);
    for (size_t t = 0; t < tiles; t++)
        for (uint32_t e = 0; e < cnt[t]; e++) {
            const size_t i = t * olap::kTileV + e;
            part[(size_t)partition_of(oid[i], size_lineorder, nthreads)].push_back({oid[i], gid[i]});
        }
    // the reference's lists are in row order within each thread's range
    for (int j = 0; j < nthreads; j++) {
        auto &v = part[(size_t)j];
        std::sort(v.begin(), v.end());
        for (size_t e = 0; e < v.size(); e++) {
            OID[j][e] = v[e].first;
            groupID[j][e] = (int16_t)v[e].second;
        }
        index[j] = (int)v.size();
    }
}

int agg_cwm_dv(int64_t **OID, int16_t **groupID, int *index, int &size_lineorder, int32_t *M1,
               int32_t *M2, int group_nums, int nthreads)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ensure_ctx()) return -1;
    if (nthreads < 1) nthreads = 1;
    std::vector<uint32_t> oid;
    std::vector<int32_t> gid;
    for (int j = 0; j < nthreads; j++)
        for (int e = 0; e < index[j]; e++) {
            oid.push_back((uint32_t)OID[j][e]);
            gid.push_back(groupID[j][e]);
        }
    DevBuf doid, dgid, dm1, dm2, sum, cnt;
    const size_t rows = (size_t)size_lineorder;
    if (!upload(doid, oid.data(), oid.size() * 4) || !upload(dgid, gid.data(), gid.size() * 4) ||
        !upload(dm1, M1, rows * 4) || (M2 && !upload(dm2, M2, rows * 4)))
        return -1;
    if (olap_dev_alloc(g_ctx, 8 * (size_t)group_nums, &sum.p) != OLAP_OK ||
        olap_dev_alloc(g_ctx, 8 * (size_t)group_nums, &cnt.p) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return -1;
    }
    if (olap::agg_row_list(g_ctx, (const uint32_t *)doid.p, (const int32_t *)dgid.p,
                           (int64_t)oid.size(), (const int32_t *)dm1.p, (const int32_t *)dm2.p,
                           group_nums, (int64_t *)sum.p, (int64_t *)cnt.p, nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return -1;
    }
    if (!fetch_results(group_nums, sum, cnt)) return -1;
    return legacy(g_sum);
}

void join_cwm_sv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                 int16_t *groupID, int *factor, int &size_lineorder, int nthreads)
{
    (void)nthreads;
    std::lock_guard<std::mutex> lk(g_mu);
    Staged S;
    if (!stage(S, dimvec_array, one_shard(fk_array, size_lineorder, nullptr, nullptr), orders,
               dimvec_nums, factor, size_lineorder, INT32_MAX))
        return;
    DevBuf g;
    if (olap_dev_alloc(g_ctx, 4 * (size_t)size_lineorder + 16, &g.p) != OLAP_OK ||
        olap_join_cwm_sv(S.q, (int32_t *)g.p, nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return;
    }
    std::vector<int32_t> h((size_t)size_lineorder);
    if (olap_copy_d2h(g_ctx, h.data(), g.p, 4 * h.size(), nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return;
    }
    for (size_t i = 0; i < h.size(); i++)
        groupID[i] = h[i] < 0 ? kGroupNull16 : (int16_t)h[i];
}

int agg_cwm_sv(int16_t *groupID, int &size_lineorder, int32_t *M1, int32_t *M2, int group_nums,
               int nthreads)
{
    (void)nthreads;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ensure_ctx()) return -1;
    const size_t rows = (size_t)size_lineorder;
    std::vector<int32_t> g(rows);
    for (size_t i = 0; i < rows; i++) g[i] = groupID[i] == kGroupNull16 ? OLAP_GID_NULL : groupID[i];
    DevBuf dg, dm1, dm2, sum, cnt;
    if (!upload(dg, g.data(), rows * 4) || !upload(dm1, M1, rows * 4) ||
        (M2 && !upload(dm2, M2, rows * 4)))
        return -1;
    if (olap_dev_alloc(g_ctx, 8 * (size_t)group_nums, &sum.p) != OLAP_OK ||
        olap_dev_alloc(g_ctx, 8 * (size_t)group_nums, &cnt.p) != OLAP_OK ||
        olap::agg_gid_vector(g_ctx, (const int32_t *)dg.p, (int64_t)rows, (const int32_t *)dm1.p,
                             (const int32_t *)dm2.p, group_nums, (int64_t *)sum.p,
                             (int64_t *)cnt.p, nullptr) != OLAP_OK) {
        g_err = olap_last_error(g_ctx);
        return -1;
    }
    if (!fetch_results(group_nums, sum, cnt)) return -1;
    return legacy(g_sum);
}
