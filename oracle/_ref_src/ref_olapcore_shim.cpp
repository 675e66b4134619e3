// Recipe piece for oracle/_ref: compiles the reference OLAPcore operator from
// its own source file, where it lies, and exposes its single-thread workers
// through a C ABI for the checker. Nothing of the reference is copied here;
// the include path is given by oracle/Makefile (REF=/root/reference).
#define main ref_olapcore_main
#include "code/OLAPcore/test_multi_compute_operator.cpp"
#undef main

extern "C" {

// gen_data, include/gendata_util.hpp:449-542, copied out of its new[] buffers
int ref_gen_data(double c_sele, double s_sele, double p_sele, double d_sele, double sf,
                 const int *bitmap, const int *groups, int8_t **dimvec_out,
                 int32_t **fk_out, int32_t *m1_out, int32_t *m2_out)
{
    int8_t *dv[4];
    int32_t *fk[4];
    int32_t *M1, *M2;
    int bm[4] = {bitmap[0], bitmap[1], bitmap[2], bitmap[3]};
    int gr[4] = {groups[0], groups[1], groups[2], groups[3]};
    gen_data(c_sele, s_sele, p_sele, d_sele, sf, bm[0], bm[1], bm[2], bm[3],
             gr[0], gr[1], gr[2], gr[3], dv[0], dv[1], dv[2], dv[3],
             fk[0], fk[1], fk[2], fk[3], M1, M2);
    const TABLE_NAME tabs[4] = {customer, supplier, part, date};
    for (int t = 0; t < 4; t++) {
        int n = size_of_table(tabs[t], sf);
        memcpy(dimvec_out[t], dv[t], (size_t)n);
        delete[] dv[t];
    }
    int n = size_of_table(lineorder, sf);
    for (int t = 0; t < 4; t++) {
        memcpy(fk_out[t], fk[t], sizeof(int32_t) * (size_t)n);
        delete[] fk[t];
    }
    memcpy(m1_out, M1, sizeof(int32_t) * (size_t)n);
    memcpy(m2_out, M2, sizeof(int32_t) * (size_t)n);
    delete[] M1;
    delete[] M2;
    return n;
}

int ref_size_of_table(int t, double sf)
{
    const TABLE_NAME tabs[5] = {customer, supplier, part, date, lineorder};
    return size_of_table(tabs[t], sf);
}

void ref_srand(unsigned s) { srand(s); }

void ref_sort_by_rate(double *sele, int *orders, int n) { sort_by_rate(sele, orders, n); }

// OLAPcore_row_thread, one worker over [0, n)
void ref_row(int8_t **dimvec, int32_t **fk, int *orders, int ndims, int *factor, int n,
             int32_t *m1, int32_t *m2, uint32_t *group_vector)
{
    pth_rowolapcoret a;
    memset(&a, 0, sizeof(a));
    a.start = 0;
    a.num_tuples = n;
    a.dimvec_array = dimvec;
    a.fk_array = fk;
    a.dimvec_nums = ndims;
    a.orders = orders;
    a.group_vector = group_vector;
    a.M1 = m1;
    a.M2 = m2;
    a.factor = factor;
    OLAPcore_row_thread(&a);
}

static void vwm_args(pth_vwmolapcoret &a, int8_t **dimvec, int32_t **fk, int *orders, int ndims,
                     int *factor, int n, int32_t *m1, int32_t *m2, uint32_t *gv,
                     int64_t *oid, int16_t *gid, int *index)
{
    memset(&a, 0, sizeof(a));
    a.start = 0;
    a.num_tuples = n;
    a.dimvec_array = dimvec;
    a.fk_array = fk;
    a.dimvec_nums = ndims;
    a.orders = orders;
    a.OID = oid;
    a.groupID = gid;
    a.group_vector = gv;
    a.M1 = m1;
    a.M2 = m2;
    a.index = index;
    a.factor = factor;
}

// OLAPcore_vwm_dv_thread, one worker
void ref_vwm_dv(int8_t **dimvec, int32_t **fk, int *orders, int ndims, int *factor, int n,
                int32_t *m1, int32_t *m2, uint32_t *group_vector)
{
    pth_vwmolapcoret a;
    int64_t oid[size_v];
    int16_t gid[size_v];
    int index = 0;
    vwm_args(a, dimvec, fk, orders, ndims, factor, n, m1, m2, group_vector, oid, gid, &index);
    OLAPcore_vwm_dv_thread(&a);
}

// OLAPcore_vwm_sv_thread, one worker
void ref_vwm_sv(int8_t **dimvec, int32_t **fk, int *orders, int ndims, int *factor, int n,
                int32_t *m1, int32_t *m2, uint32_t *group_vector)
{
    pth_vwmolapcoret a;
    int64_t oid[size_v];
    int16_t gid[size_v];
    int index = 0;
    vwm_args(a, dimvec, fk, orders, ndims, factor, n, m1, m2, group_vector, oid, gid, &index);
    OLAPcore_vwm_sv_thread(&a);
}

// join_cwm_dv_thread per probe step, then agg_cwm_dv_thread, one worker
void ref_cwm_dv(int8_t **dimvec, int32_t **fk, int *orders, int ndims, int *factor, int n,
                int32_t *m1, int32_t *m2, uint32_t *group_vector)
{
    int64_t *oid = new int64_t[n > 0 ? n : 1];
    int16_t *gid = new int16_t[n > 0 ? n : 1];
    int index = 0;
    for (int i = 0; i < ndims; i++) {
        pth_cwmjoint j;
        memset(&j, 0, sizeof(j));
        j.join_id = i;
        j.start = 0;
        j.num_tuples = n;
        j.dimvec = dimvec[orders[i]];
        j.fk = fk[orders[i]];
        j.OID = oid;
        j.groupID = gid;
        j.index = &index;
        j.factor = factor[i];
        join_cwm_dv_thread(&j);
    }
    pth_cwmaggt g;
    memset(&g, 0, sizeof(g));
    g.start = 0;
    g.num_tuples = n;
    g.OID = oid;
    g.groupID = gid;
    g.index = &index;
    g.M1 = m1;
    g.M2 = m2;
    g.group_vector = group_vector;
    agg_cwm_dv_thread(&g);
    delete[] oid;
    delete[] gid;
}

// join_cwm_sv_thread per probe step, then agg_cwm_sv_thread, one worker;
// gid_out receives the int16 static vector
void ref_cwm_sv(int8_t **dimvec, int32_t **fk, int *orders, int ndims, int *factor, int n,
                int32_t *m1, int32_t *m2, uint32_t *group_vector, int16_t *gid_out)
{
    memset(gid_out, 0, sizeof(int16_t) * (size_t)n);
    for (int i = 0; i < ndims; i++) {
        pth_cwmjoint j;
        memset(&j, 0, sizeof(j));
        j.join_id = i;
        j.start = 0;
        j.num_tuples = n;
        j.dimvec = dimvec[orders[i]];
        j.fk = fk[orders[i]];
        j.groupID = gid_out;
        j.factor = factor[i];
        join_cwm_sv_thread(&j);
    }
    pth_cwmaggt g;
    memset(&g, 0, sizeof(g));
    g.start = 0;
    g.num_tuples = n;
    g.groupID = gid_out;
    g.M1 = m1;
    g.M2 = m2;
    g.group_vector = group_vector;
    agg_cwm_sv_thread(&g);
}

// the multi-threaded driver, OLAPcore_vwm_dv :1052-1126, used as the timed CPU
// baseline; it pins worker j to CPU j
int ref_OLAPcore_vwm_dv(int8_t **dimvec, int32_t **fk, int *orders, int ndims, int *factor,
                        int n, int group_nums, int32_t *m1, int32_t *m2, int nthreads)
{
    return OLAPcore_vwm_dv(dimvec, fk, orders, ndims, factor, n, group_nums, m1, m2, nthreads);
}

int ref_OLAPcore_rowwise(int8_t **dimvec, int32_t **fk, int *orders, int ndims, int *factor,
                         int n, int group_nums, int32_t *m1, int32_t *m2, int nthreads)
{
    return OLAPcore_rowwise(dimvec, fk, orders, ndims, factor, n, group_nums, m1, m2, nthreads);
}

}  // extern "C"
