// Drives the reference-signature layer (include/olapcore_compat.hpp) the way
// the reference's test_multi_compute_operator.cpp calls its operators, on a
// star read from raw files, and writes what every call returned.
//
//   compat_harness <dir>
//
// <dir>/meta.txt: rows ndims G nthreads numa_nodes, then orders[ndims],
// factor[ndims], dim_rows[4]; dimvec<t>.bin (int8, dim_rows[t]), fk<t>.bin
// (int32, rows), m1.bin, m2.bin (int32, rows). Writes <dir>/out.txt (one
// "name value" line per call's returned int), sums_<name>.bin and
// counts_<name>.bin (int64 x G, the exact results of the call),
// cwm_dv_index.bin (int32 x nthreads), cwm_dv_oid<j>.bin / cwm_dv_gid<j>.bin
// (int64 / int16 per thread list) and cwm_sv_gid.bin (int16 x rows).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "olapcore_compat.hpp"

// the reference's NUMA argument structs, as include/metadata.h:272-279 has them
struct Dimvec_array_numa {
    int8_t *dimvec[4];
};
struct Fk_array_numa {
    int32_t *fk[4];
};

namespace {

std::string dir;

template <class T>
std::vector<T> load(const std::string &name, size_t n)
{
    std::vector<T> v(n ? n : 1);
    FILE *f = fopen((dir + "/" + name).c_str(), "rb");
    if (!f || fread(v.data(), sizeof(T), n, f) != n) {
        fprintf(stderr, "cannot read %s\n", name.c_str());
        exit(2);
    }
    fclose(f);
    return v;
}

template <class T>
void save(const std::string &name, const T *p, size_t n)
{
    FILE *f = fopen((dir + "/" + name).c_str(), "wb");
    if (!f || (n && fwrite(p, sizeof(T), n, f) != n)) {
        fprintf(stderr, "cannot write %s\n", name.c_str());
        exit(2);
    }
    fclose(f);
}

FILE *out;
int G;

void record(const char *name, int ret)
{
    if (ret == -1) fprintf(stderr, "%s: %s\n", name, olapcore_compat_last_error());
    fprintf(out, "%s %d\n", name, ret);
    save(std::string("sums_") + name + ".bin", olapcore_compat_last_sums(), (size_t)G);
    save(std::string("counts_") + name + ".bin", olapcore_compat_last_counts(), (size_t)G);
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: compat_harness <dir>\n");
        return 2;
    }
    dir = argv[1];
    FILE *m = fopen((dir + "/meta.txt").c_str(), "r");
    int rows = 0, ndims = 0, nthreads = 1, nodes = 1;
    int orders[4] = {0, 0, 0, 0}, factor[4] = {0, 0, 0, 0};
    long long dr[4];
    if (!m || fscanf(m, "%d %d %d %d %d", &rows, &ndims, &G, &nthreads, &nodes) != 5) return 2;
    for (int i = 0; i < ndims; i++)
        if (fscanf(m, "%d", &orders[i]) != 1) return 2;
    for (int i = 0; i < ndims; i++)
        if (fscanf(m, "%d", &factor[i]) != 1) return 2;
    for (int t = 0; t < 4; t++)
        if (fscanf(m, "%lld", &dr[t]) != 1) return 2;
    fclose(m);

    std::vector<int8_t> dv[4];
    std::vector<int32_t> fk[4];
    int8_t *dimvec_array[4];
    int32_t *fk_array[4];
    int64_t dim_rows[4];
    for (int t = 0; t < 4; t++) {
        dv[t] = load<int8_t>("dimvec" + std::to_string(t) + ".bin", (size_t)dr[t]);
        fk[t] = load<int32_t>("fk" + std::to_string(t) + ".bin", (size_t)rows);
        dimvec_array[t] = dv[t].data();
        fk_array[t] = fk[t].data();
        dim_rows[t] = dr[t];
    }
    std::vector<int32_t> M1 = load<int32_t>("m1.bin", (size_t)rows);
    std::vector<int32_t> M2 = load<int32_t>("m2.bin", (size_t)rows);
    olapcore_compat_set_dim_rows(dim_rows);
    out = fopen((dir + "/out.txt").c_str(), "w");
    if (!out) return 2;

    // fused drivers, called as test_OLAPcore_rowwise / _vectorwise call them
    record("rowwise", OLAPcore_rowwise(dimvec_array, fk_array, orders, ndims, factor, rows, G,
                                       M1.data(), M2.data(), nthreads));
    record("vwm_dv", OLAPcore_vwm_dv(dimvec_array, fk_array, orders, ndims, factor, rows, G,
                                     M1.data(), M2.data(), nthreads));
    record("vwm_sv", OLAPcore_vwm_sv(dimvec_array, fk_array, orders, ndims, factor, rows, G,
                                     M1.data(), M2.data(), nthreads));

    // column-wise dynamic vector: per-thread lists sized to the whole fact
    // table, as test_OLAPcore_cwm_dv allocates them (:1504-1510)
    std::vector<std::vector<int64_t>> oid((size_t)nthreads, std::vector<int64_t>((size_t)rows + 1));
    std::vector<std::vector<int16_t>> gid((size_t)nthreads, std::vector<int16_t>((size_t)rows + 1));
    std::vector<int64_t *> OID((size_t)nthreads);
    std::vector<int16_t *> groupID((size_t)nthreads);
    for (int j = 0; j < nthreads; j++) {
        OID[(size_t)j] = oid[(size_t)j].data();
        groupID[(size_t)j] = gid[(size_t)j].data();
    }
    std::vector<int> index((size_t)nthreads, -1);
    join_cwm_dv(dimvec_array, fk_array, orders, ndims, OID.data(), groupID.data(), factor,
                index.data(), rows, nthreads);
    save("cwm_dv_index.bin", index.data(), (size_t)nthreads);
    for (int j = 0; j < nthreads; j++) {
        save("cwm_dv_oid" + std::to_string(j) + ".bin", OID[(size_t)j], (size_t)index[(size_t)j]);
        save("cwm_dv_gid" + std::to_string(j) + ".bin", groupID[(size_t)j], (size_t)index[(size_t)j]);
    }
    record("cwm_dv", agg_cwm_dv(OID.data(), groupID.data(), index.data(), rows, M1.data(), M2.data(),
                                G, nthreads));

    // column-wise static vector: one int16 gid per fact row
    std::vector<int16_t> sv((size_t)rows + 1, 0);
    join_cwm_sv(dimvec_array, fk_array, orders, ndims, sv.data(), factor, rows, nthreads);
    save("cwm_sv_gid.bin", sv.data(), (size_t)rows);
    record("cwm_sv", agg_cwm_sv(sv.data(), rows, M1.data(), M2.data(), G, nthreads));

    // NUMA drivers: node shards of the fact columns as the reference's main
    // cuts them (:2017-2057), replicas of the dimension vectors
    olapcore_compat_set_numa_nodes(nodes);
    std::vector<Dimvec_array_numa> dvn((size_t)nodes);
    std::vector<Fk_array_numa> fkn((size_t)nodes);
    std::vector<int32_t *> m1n((size_t)nodes), m2n((size_t)nodes);
    std::vector<std::vector<int8_t>> rep((size_t)nodes * 4);
    std::vector<std::vector<int32_t>> shard((size_t)nodes * 6);
    for (int i = 0; i < nodes; i++) {
        const int per = rows / nodes;
        const int n = i == nodes - 1 ? rows - (nodes - 1) * per : per;
        const size_t first = (size_t)i * (size_t)per;
        for (int t = 0; t < 4; t++) {
            rep[(size_t)i * 4 + t] = dv[t];
            dvn[(size_t)i].dimvec[t] = rep[(size_t)i * 4 + t].data();
            auto &s = shard[(size_t)i * 6 + t];
            s.assign(fk[t].begin() + first, fk[t].begin() + first + n);
            s.push_back(0);
            fkn[(size_t)i].fk[t] = s.data();
        }
        auto &a = shard[(size_t)i * 6 + 4];
        auto &b = shard[(size_t)i * 6 + 5];
        a.assign(M1.begin() + first, M1.begin() + first + n);
        b.assign(M2.begin() + first, M2.begin() + first + n);
        a.push_back(0);
        b.push_back(0);
        m1n[(size_t)i] = a.data();
        m2n[(size_t)i] = b.data();
    }
    record("vwm_dv_numa", OLAPcore_vwm_dv_numa(dvn.data(), fkn.data(), orders, ndims, factor, rows,
                                               G, m1n.data(), m2n.data(), nthreads));
    record("vwm_sv_numa", OLAPcore_vwm_sv_numa(dvn.data(), fkn.data(), orders, ndims, factor, rows,
                                               G, m1n.data(), m2n.data(), nthreads));
    fclose(out);
    return 0;
}
