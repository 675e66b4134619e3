// olapcore_compat.hpp -- the reference's OLAPcore operator functions, same
// names, same C++ signatures, same argument meaning, implemented on the GPU
// through the C ABI (olapcore.h). A caller of the reference operator
// (code/OLAPcore/test_multi_compute_operator.cpp) links libolapcore.so in
// place of those functions.
//
// Differences a caller can observe, all deliberate:
//  - columns are host arrays as in the reference; every call uploads them to
//    the GPU (the resident, zero-copy path is the C ABI);
//  - the returned int is the reference's: the uint32 group vector summed into
//    an int (test_multi_compute_operator.cpp:1027-1036), computed from exact
//    int64 sums, so it wraps exactly where the reference's wraps;
//  - a foreign key outside [0, dimension rows) finds no row (the reference
//    reads past the vector); the dimension sizes are therefore required and
//    are set once with olapcore_compat_set_dim_rows;
//  - nthreads is accepted and ignored (one GPU); the NUMA drivers read the
//    node shards in row order and run one scan over all of them;
//  - errors return -1 (message in olapcore_compat_last_error) instead of exit(-1).
#pragma once

#include <stdint.h>

// Dimension-vector lengths by table slot (0 customer, 1 supplier, 2 part,
// 3 date), e.g. size_of_table() of the reference's generator.
void olapcore_compat_set_dim_rows(const int64_t rows[4]);
// Node shards the NUMA drivers' arrays hold (default 0: the host's NUMA
// nodes, numa_max_node() + 1, as the reference's main sizes them).
void olapcore_compat_set_numa_nodes(int nodes);
// Device used by the compat functions (default 0).
void olapcore_compat_set_device(int device);
const char *olapcore_compat_last_error();
// int64 (sum, count) of the last call, group_nums entries each (host).
const int64_t *olapcore_compat_last_sums();
const int64_t *olapcore_compat_last_counts();

// OLAPcore_rowwise, test_multi_compute_operator.cpp:973
int OLAPcore_rowwise(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                     int *factor, int &size_lineorder, int &group_nums, int32_t *M1, int32_t *M2,
                     int nthreads);
// OLAPcore_vwm_dv, :1052
int OLAPcore_vwm_dv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                    int *factor, int &size_lineorder, int &group_nums, int32_t *M1, int32_t *M2,
                    int nthreads);
// OLAPcore_vwm_sv, :1258
int OLAPcore_vwm_sv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                    int *factor, int &size_lineorder, int &group_nums, int32_t *M1, int32_t *M2,
                    int nthreads);
// The reference's NUMA argument structs (include/metadata.h:272-279), defined
// there as { int8_t *dimvec[4]; } and { int32_t *fk[4]; }, one per node.
struct Dimvec_array_numa;
struct Fk_array_numa;
// OLAPcore_vwm_dv_numa, :1142 -- node i holds rows [i * (N / n), ...) of the
// fact columns (fk_array_numa[i], M1_p[i], M2_p[i]; the last node the
// remainder) and a replica of the dimension vectors (dimvec_array_numa[i])
int OLAPcore_vwm_dv_numa(Dimvec_array_numa *dimvec_array_numa, Fk_array_numa *fk_array_numa,
                         int *orders, int &dimvec_nums, int *factor, int &size_lineorder,
                         int &group_nums, int32_t **M1_p, int32_t **M2_p, int nthreads);
// OLAPcore_vwm_sv_numa, :1342
int OLAPcore_vwm_sv_numa(Dimvec_array_numa *dimvec_array_numa, Fk_array_numa *fk_array_numa,
                         int *orders, int &dimvec_nums, int *factor, int &size_lineorder,
                         int &group_nums, int32_t **M1_p, int32_t **M2_p, int nthreads);
// join_cwm_dv, :840 -- fills the per-thread OID/groupID lists and index[]
// counts exactly as the reference partitions rows (thread j owns rows
// [j*N/nthreads, ...), the last thread the remainder)
void join_cwm_dv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                 int64_t **OID, int16_t **groupID, int *factor, int *index, int &size_lineorder,
                 int nthreads);
// agg_cwm_dv, :685
int agg_cwm_dv(int64_t **OID, int16_t **groupID, int *index, int &size_lineorder, int32_t *M1,
               int32_t *M2, int group_nums, int nthreads);
// join_cwm_sv, :908 -- groupID[row] = gid or GROUP_NULL (32767)
void join_cwm_sv(int8_t **dimvec_array, int32_t **fk_array, int *orders, int &dimvec_nums,
                 int16_t *groupID, int *factor, int &size_lineorder, int nthreads);
// agg_cwm_sv, :764
int agg_cwm_sv(int16_t *groupID, int &size_lineorder, int32_t *M1, int32_t *M2, int group_nums,
               int nthreads);
