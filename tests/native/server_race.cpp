// Stress of the serving loop's host side under ThreadSanitizer: a server with
// a host executor (no device), four submitting threads, waits with and without
// timeouts, cancellations, pause/resume and statistics reads racing with the
// workers. Every completed request is checked against its closed form: the
// executor fills slot g of a query with (dataset + 1) * 1000 + g and count g.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "olapcore.h"

namespace {

olap_status exec_fn(void *, int32_t dataset, const olap_batch_query *qs, int32_t n, const int64_t *off,
                    int64_t *sum, int64_t *count)
{
    for (int32_t i = 0; i < n; i++)
        for (int64_t g = 0; g < qs[i].group_nums; g++) {
            sum[off[i] + g] = (int64_t)(dataset + 1) * 1000 + g;
            count[off[i] + g] = g;
        }
    return OLAP_OK;
}

}  // namespace

int main()
{
    olap_server_config cfg{16, 3, 1 << 20};
    olap_server s = nullptr;
    if (olap_server_create_host(&cfg, exec_fn, nullptr, &s) != OLAP_OK) return 2;
    static int32_t fake[4];  // dataset columns are never read by a host executor
    olap_star star;
    std::memset(&star, 0, sizeof(star));
    for (int t = 0; t < 4; t++) {
        star.filter[t] = fake;
        star.attr[t] = fake;
        star.fk[t] = fake;
        star.dim_rows[t] = 64;
    }
    int32_t ids[2];
    if (olap_server_add_dataset(s, &star, &ids[0]) != OLAP_OK || olap_server_add_dataset(s, &star, &ids[1]) != OLAP_OK)
        return 3;
    std::atomic<int> bad{0}, done{0}, cancelled{0};
    auto client = [&](int k) {
        for (int i = 0; i < 400; i++) {
            olap_batch_query q;
            std::memset(&q, 0, sizeof(q));
            q.ndims = 1;
            q.table[0] = (k + i) & 3;
            q.group_mod[0] = 1 + (i % 7);
            q.factor[0] = 1;
            q.group_nums = 1 + (i % 7);
            std::vector<int64_t> sum((size_t)q.group_nums), cnt((size_t)q.group_nums);
            const int d = (k + i) & 1;
            uint64_t t = 0;
            if (olap_server_submit(s, ids[d], &q, sum.data(), cnt.data(), &t) != OLAP_OK) {
                bad++;
                continue;
            }
            if (i % 9 == 0 && olap_server_cancel(s, t) == OLAP_OK) {
                cancelled++;
                continue;
            }
            int32_t st = -1;
            double lat = 0;
            olap_status w;
            while ((w = olap_server_wait(s, t, i % 2 ? 50 : -1, &st, &lat)) == OLAP_E_BUSY) {
            }
            if (w != OLAP_OK || st != OLAP_OK) {
                bad++;
                continue;
            }
            for (int64_t g = 0; g < q.group_nums; g++)
                if (sum[(size_t)g] != (int64_t)(d + 1) * 1000 + g || cnt[(size_t)g] != g) bad++;
            done++;
        }
    };
    std::vector<std::thread> th;
    for (int k = 0; k < 4; k++) th.emplace_back(client, k);
    for (int r = 0; r < 50; r++) {
        olap_server_pause(s, r & 1);
        olap_server_stats stt;
        olap_server_get_stats(s, &stt);
        std::this_thread::yield();
    }
    olap_server_pause(s, 0);
    for (auto &t : th) t.join();
    olap_server_stats stt;
    olap_server_get_stats(s, &stt);
    olap_server_destroy(s);
    std::printf("completed %d cancelled %d bad %d batches %lld\n", done.load(), cancelled.load(), bad.load(),
                (long long)stt.batches);
    return bad.load() == 0 && done.load() + cancelled.load() == 1600 ? 0 : 1;
}
