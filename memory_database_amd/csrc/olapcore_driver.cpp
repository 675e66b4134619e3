// olapcore_driver.cpp -- command-line driver of the GPU star-join group-by,
// with the reference's flags (code/OLAPcore/test_multi_compute_operator.cpp:
// 36-60, :76-100): --SF --{c,s,p,d}-sele --{c,s,p,d}-bitmap --{c,s,p,d}-groups
// --nthreads --sqlnum --basic_numa. Extensions: --strategy --gen --pad --seed
// --device --repeat --log --json.
//
// --gen=device (default): dimension vectors and fact columns are generated on
//   the GPU with the counter-based generator (FKs uniform over [0, D)).
// --gen=reference: the reference's gen_data (include/gendata_util.hpp:449-542)
//   restated on the host with glibc rand(): same draw order, FKs over [0, D]
//   inclusive. The byte past each dimension vector that the reference reads
//   for FK == D is made explicit with --pad (127 = NULL, 0 = group 0).
//
// Output follows the reference's lines ("Result count of total selection rate
// ... is C/N", "total time: T ms") and appends its TSV row format.
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <chrono>
#include <fstream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "olapcore.h"

namespace {

struct Args {
    double sf = 1.0;
    double sele[4] = {0, 0, 0, 0};  // c, s, p, d
    int bitmap[4] = {0, 0, 0, 0};
    int groups[4] = {0, 0, 0, 0};
    int nthreads = 2;
    int sqlnum = 0;
    int basic_numa = 0;
    std::string strategy = "all";
    std::string gen = "device";
    int pad = 127;
    unsigned long long seed = 20251203ull;
    int device = 0;
    int repeat = 1;
    std::string log = "log/multi_compute_operator/olapcore_test.tsv";
    std::string json;
};

// size_of_table, include/gendata_util.hpp:152-166 (bases metadata.h:71-75)
int64_t table_rows(int t, double sf)
{
    switch (t) {
    case 0: return (int64_t)(int)(30000.0 * sf);
    case 1: return (int64_t)(int)(2000.0 * sf);
    case 2: return (int64_t)(int)(200000.0 * (1.0 + log2(sf)));
    case 3: return 7 * 365;
    default: return (int64_t)(int)(6000000.0 * sf);
    }
}

void usage(const char *prog)
{
    printf("Usage: %s [options]\n"
           "  --SF=<sf>                       scale factor (fact rows = 6,000,000 * SF)\n"
           "  --{c,s,p,d}-sele=<rate>         selection rate of customer/supplier/part/date\n"
           "  --{c,s,p,d}-bitmap=<0|1>        1: the dimension bears groups, 0: filter only\n"
           "  --{c,s,p,d}-groups=<n>          groups of the dimension (0: not joined)\n"
           "  --nthreads=<n> --sqlnum=<n> --basic_numa   accepted as in the reference\n"
           "  --strategy=all|rowwise|vwm_dv|vwm_sv|cwm_dv|cwm_sv\n"
           "  --gen=device|reference   --pad=<byte>   --seed=<n>   --device=<n>\n"
           "  --repeat=<n>   --log=<tsv path>   --json=<jsonl path>\n",
           prog);
}

bool parse(int argc, char **argv, Args &a)
{
    enum {
        O_SF = 1000, O_NT, O_SQL, O_NUMA, O_STRAT, O_GEN, O_PAD, O_SEED, O_DEV, O_REP, O_LOG, O_JSON,
        O_SELE = 1100, O_BITMAP = 1200, O_GROUPS = 1300
    };
    static const char *tabs = "cspd";
    std::vector<option> opts = {
        {"help", no_argument, nullptr, 'h'},        {"SF", required_argument, nullptr, O_SF},
        {"nthreads", required_argument, nullptr, O_NT}, {"sqlnum", required_argument, nullptr, O_SQL},
        {"basic_numa", no_argument, nullptr, O_NUMA}, {"strategy", required_argument, nullptr, O_STRAT},
        {"gen", required_argument, nullptr, O_GEN},  {"pad", required_argument, nullptr, O_PAD},
        {"seed", required_argument, nullptr, O_SEED}, {"device", required_argument, nullptr, O_DEV},
        {"repeat", required_argument, nullptr, O_REP}, {"log", required_argument, nullptr, O_LOG},
        {"json", required_argument, nullptr, O_JSON}};
    static char names[12][16];
    for (int t = 0; t < 4; t++) {
        snprintf(names[t], 16, "%c-sele", tabs[t]);
        snprintf(names[4 + t], 16, "%c-bitmap", tabs[t]);
        snprintf(names[8 + t], 16, "%c-groups", tabs[t]);
        opts.push_back({names[t], required_argument, nullptr, O_SELE + t});
        opts.push_back({names[4 + t], required_argument, nullptr, O_BITMAP + t});
        opts.push_back({names[8 + t], required_argument, nullptr, O_GROUPS + t});
    }
    opts.push_back({nullptr, 0, nullptr, 0});
    int c;
    while ((c = getopt_long(argc, argv, "hn:", opts.data(), nullptr)) != -1) {
        if (c == 'h' || c == '?') {
            usage(argv[0]);
            return false;
        }
        if (c == 'n' || c == O_NT) a.nthreads = atoi(optarg);
        else if (c == O_SF) a.sf = atof(optarg);
        else if (c == O_SQL) a.sqlnum = atoi(optarg);
        else if (c == O_NUMA) a.basic_numa = 1;
        else if (c == O_STRAT) a.strategy = optarg;
        else if (c == O_GEN) a.gen = optarg;
        else if (c == O_PAD) a.pad = atoi(optarg);
        else if (c == O_SEED) a.seed = strtoull(optarg, nullptr, 10);
        else if (c == O_DEV) a.device = atoi(optarg);
        else if (c == O_REP) a.repeat = atoi(optarg);
        else if (c == O_LOG) a.log = optarg;
        else if (c == O_JSON) a.json = optarg;
        else if (c >= O_SELE && c < O_SELE + 4) a.sele[c - O_SELE] = atof(optarg);
        else if (c >= O_BITMAP && c < O_BITMAP + 4) a.bitmap[c - O_BITMAP] = atoi(optarg);
        else if (c >= O_GROUPS && c < O_GROUPS + 4) a.groups[c - O_GROUPS] = atoi(optarg);
    }
    return true;
}

#define CHECK(ctx, call)                                                                    \
    do {                                                                                    \
        olap_status st_ = (call);                                                           \
        if (st_ != OLAP_OK) {                                                               \
            fprintf(stderr, "%s failed: status %d: %s\n", #call, (int)st_, olap_last_error(ctx)); \
            exit(1);                                                                        \
        }                                                                                   \
    } while (0)

void *dalloc(olap_ctx ctx, size_t bytes)
{
    void *p = nullptr;
    CHECK(ctx, olap_dev_alloc(ctx, bytes, &p));
    return p;
}

// gen_data of the reference on the host, glibc rand() in its draw order
void gen_reference(const Args &a, std::vector<std::vector<int8_t>> &dv,
                   std::vector<std::vector<int32_t>> &fk, std::vector<int32_t> &m, int64_t n,
                   const int64_t D[4])
{
    srand((unsigned)(a.seed == 20251203ull ? 1 : a.seed));  // never seeded there: glibc seed 1
    for (int t = 0; t < 4; t++) {
        dv[t].assign((size_t)D[t] + 1, (int8_t)a.pad);
        for (int64_t i = 0; i < D[t]; i++) {
            const double u = (double)rand() / (double)(RAND_MAX / 1.0);
            if (a.bitmap[t])
                dv[t][i] = u <= a.sele[t] ? (int8_t)(rand() % (a.groups[t] - 1 - 0 + 1)) : (int8_t)127;
            else
                dv[t][i] = u <= a.sele[t] ? 0 : (int8_t)127;
        }
    }
    for (int t = 0; t < 4; t++) fk[t].resize((size_t)n);
    m.assign((size_t)n, 5);
    for (int64_t i = 0; i < n; i++)
        for (int t = 0; t < 4; t++) fk[t][i] = rand() % ((int)D[t] + 1);
}

const char *model_name(olap_strategy s)
{
    switch (s) {
    case OLAP_ROWWISE: return "Row-wise query model";
    case OLAP_CWM_DV: return "Column-wise query model with dynamic vector";
    case OLAP_CWM_SV: return "Column-wise query model with static vector";
    case OLAP_VWM_DV: return "Vector-wise query model with dynamic vector";
    default: return "Vector-wise query model with static vector";
    }
}

}  // namespace

int main(int argc, char **argv)
{
    Args a;
    if (!parse(argc, argv, a)) return 0;
    olap_ctx ctx = nullptr;
    olap_status st = olap_ctx_create(a.device, &ctx);
    if (st != OLAP_OK) {
        fprintf(stderr, "no usable GPU (status %d)\n", (int)st);
        return 2;
    }
    int64_t D[4];
    for (int t = 0; t < 4; t++) D[t] = table_rows(t, a.sf);
    const int64_t n = table_rows(4, a.sf);
    olap_plan plan;
    CHECK(ctx, olap_make_plan(a.groups, a.bitmap, a.sele, &plan));

    // columns in HBM
    olap_columns cols;
    memset(&cols, 0, sizeof(cols));
    std::vector<void *> owned;
    auto t0 = std::chrono::steady_clock::now();
    if (a.gen == "reference") {
        std::vector<std::vector<int8_t>> dv(4);
        std::vector<std::vector<int32_t>> fk(4);
        std::vector<int32_t> m;
        gen_reference(a, dv, fk, m, n, D);
        for (int t = 0; t < 4; t++) {
            void *d = dalloc(ctx, dv[t].size());
            CHECK(ctx, olap_copy_h2d(ctx, d, dv[t].data(), dv[t].size(), nullptr));
            void *f = dalloc(ctx, 4 * (size_t)n);
            CHECK(ctx, olap_copy_h2d(ctx, f, fk[t].data(), 4 * (size_t)n, nullptr));
            cols.dimvec[t] = d;
            cols.dim_rows[t] = D[t] + 1;  // the reference's inclusive key range
            cols.fk[t] = (const int32_t *)f;
            owned.push_back(d);
            owned.push_back(f);
        }
        void *mm = dalloc(ctx, 4 * (size_t)n);
        CHECK(ctx, olap_copy_h2d(ctx, mm, m.data(), 4 * (size_t)n, nullptr));
        cols.m1 = cols.m2 = (const int32_t *)mm;  // M1 = M2 = 5
        owned.push_back(mm);
    } else {
        for (int t = 0; t < 4; t++) {
            if (!a.groups[t]) continue;
            int32_t *filter = (int32_t *)dalloc(ctx, 4 * (size_t)D[t]);
            int32_t *group = (int32_t *)dalloc(ctx, 4 * (size_t)D[t]);
            int8_t *dv = (int8_t *)dalloc(ctx, (size_t)D[t]);
            int32_t *fk = (int32_t *)dalloc(ctx, 4 * (size_t)n);
            CHECK(ctx, olap_gen_uniform_i32(ctx, filter, D[t], 0, a.seed, 100 + t, 0, 1000000, nullptr));
            CHECK(ctx, olap_gen_uniform_i32(ctx, group, D[t], 0, a.seed, 200 + t, 0, a.groups[t], nullptr));
            const int32_t hi = (int32_t)llround(a.sele[t] * 1000000.0);
            CHECK(ctx, olap_build_dimvec(ctx, filter, group, D[t], INT32_MIN, hi, a.bitmap[t],
                                         a.bitmap[t] ? a.groups[t] : 0, 1, dv, nullptr));
            CHECK(ctx, olap_gen_uniform_i32(ctx, fk, n, 0, a.seed, 300 + t, 0, (int32_t)D[t], nullptr));
            cols.dimvec[t] = dv;
            cols.dim_rows[t] = D[t];
            cols.fk[t] = fk;
            owned.insert(owned.end(), {filter, group, dv, fk});
        }
        int32_t *m = (int32_t *)dalloc(ctx, 4 * (size_t)n);
        CHECK(ctx, olap_gen_uniform_i32(ctx, m, n, 0, a.seed, 400, 5, 6, nullptr));  // M = 5
        cols.m1 = cols.m2 = m;
        owned.push_back(m);
    }
    for (int t = 0; t < 4; t++) cols.code_bytes[t] = 1;
    cols.rows = n;
    CHECK(ctx, olap_stream_sync(ctx, nullptr));
    const double gen_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    printf(">>> Generated data %lld lines in lineorder table. (%s generator, %.0f ms)\n", (long long)n,
           a.gen.c_str(), gen_ms);

    double total_rate = 1;
    for (int i = 0; i < plan.dimvec_nums; i++) total_rate *= a.sele[plan.orders[i]];
    std::vector<olap_strategy> strats;
    const char *names[] = {"rowwise", "vwm_dv", "vwm_sv", "cwm_dv", "cwm_sv"};
    if (a.strategy == "all") strats = {OLAP_ROWWISE, OLAP_CWM_DV, OLAP_CWM_SV, OLAP_VWM_DV, OLAP_VWM_SV};
    else
        for (int s = 0; s < 5; s++)
            if (a.strategy == names[s]) strats.push_back((olap_strategy)s);
    if (strats.empty()) {
        fprintf(stderr, "unknown strategy %s\n", a.strategy.c_str());
        return 1;
    }
    // reference log: truncate with a header on sqlnum 21, append otherwise (:1911-1926)
    std::ofstream tsv;
    if (!a.log.empty()) {
        std::string dir = a.log.substr(0, a.log.find_last_of('/'));
        if (dir != a.log) {
            std::string cur;
            for (size_t i = 0; i <= dir.size(); i++)
                if (i == dir.size() || dir[i] == '/') {
                    cur = dir.substr(0, i);
                    if (!cur.empty()) mkdir(cur.c_str(), 0755);
                }
        }
        tsv.open(a.log, a.sqlnum == 21 ? std::ios::out | std::ios::trunc : std::ios::app);
        if (a.sqlnum == 21)
            tsv << "query number\tquery model\tintermediate results\tquery model with different "
                   "intermediate result\tphase type\truntimes(ms)\n";
    }
    FILE *js = a.json.empty() ? nullptr : fopen(a.json.c_str(), "a");

    olap_query q = nullptr;
    CHECK(ctx, olap_query_create(ctx, &plan, &cols, &q));
    const int64_t G = plan.group_nums;
    int64_t *dsum = (int64_t *)dalloc(ctx, 8 * (size_t)G);
    int64_t *dcnt = (int64_t *)dalloc(ctx, 8 * (size_t)G);
    std::vector<int64_t> hsum((size_t)G), hcnt((size_t)G);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (olap_strategy s : strats) {
        printf(">>> Start test GPU OLAPcore using %s\n", model_name(s));
        CHECK(ctx, olap_query_run(q, s, dsum, dcnt, nullptr));  // warm
        double best = 1e30;
        for (int r = 0; r < std::max(1, a.repeat); r++) {
            (void)hipEventRecord(e0, nullptr);
            CHECK(ctx, olap_query_run(q, s, dsum, dcnt, nullptr));
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            best = std::min(best, (double)ms);
        }
        CHECK(ctx, olap_copy_d2h(ctx, hsum.data(), dsum, 8 * (size_t)G, nullptr));
        CHECK(ctx, olap_copy_d2h(ctx, hcnt.data(), dcnt, 8 * (size_t)G, nullptr));
        const int32_t count = olap_legacy_count(hsum.data(), G);
        int64_t rows_pass = 0;
        for (int64_t g = 0; g < G; g++) rows_pass += hcnt[(size_t)g];
        printf("          Result count of total selection rate %g%% is %d/%lld\n", total_rate * 100,
               count, (long long)n);
        printf("          total time: %gms  (%lld passing rows, %.1f G rows/s)\n", best,
               (long long)rows_pass, n / best / 1e6);
        if (tsv.is_open()) {
            const char *inter = s == OLAP_ROWWISE ? "" : (s == OLAP_CWM_DV || s == OLAP_VWM_DV ? "dynamic vector" : "static vector");
            const char *model = s == OLAP_ROWWISE ? "Row-wise query model" : (s == OLAP_CWM_DV || s == OLAP_CWM_SV ? "Column-wise query model" : "Vector-wise query model");
            tsv << a.sqlnum << "\t" << model << "\t" << inter << "\t" << model_name(s) << "\ttotal\t" << best << "\n";
        }
        if (js)
            fprintf(js, "{\"sqlnum\": %d, \"strategy\": \"%s\", \"SF\": %g, \"rows\": %lld, "
                        "\"ms\": %.6f, \"count\": %d, \"passing_rows\": %lld, \"gen\": \"%s\"}\n",
                    a.sqlnum, names[(int)s], a.sf, (long long)n, best, count, (long long)rows_pass,
                    a.gen.c_str());
    }
    if (js) fclose(js);
    olap_query_destroy(q);
    olap_dev_free(ctx, dsum);
    olap_dev_free(ctx, dcnt);
    for (void *p : owned) olap_dev_free(ctx, p);
    olap_ctx_destroy(ctx);
    return 0;
}
