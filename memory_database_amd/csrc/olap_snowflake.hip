// olap_snowflake.hip -- TPC-H Q5's fused snowflake probe on gfx950.
//
// Per fact (lineitem) row: the supplier vector gives the supplier's nation
// code for the query's region (NULL otherwise), the orders vector gives the
// ordering customer's nation code for the region and date range; the row
// counts when both are non-NULL and equal (customer and supplier in the same
// nation), and its revenue price x (100 - discount%) goes to that nation's
// int64 slot. Both keys stream through 128-bit loads and both vectors are
// probed at once; price and discount are read only for 4-row chunks with a
// matching row;
// the per-nation table (up to kSnowSlots slots) sits in LDS.
#include "olap_internal.h"

namespace olap {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kSnowBS = 512;
constexpr int kSnowSlots = 1024;  // largest group count (OLAP_TRI_MAX_GROUPS)
constexpr int kSnowU = 4;  // 128-bit chunks per lane per tile

struct SnowArgs {
    const int32_t *skey, *okey, *price, *disc;
    const void *svec, *ovec;
    int64_t srows, orows;
    int32_t sbias, obias, sbytes, obytes;
    int64_t rows;
    int32_t G;
    unsigned long long *sum, *cnt;
    uint32_t *bad;  // rows whose code is outside [0, G)
};

template <int BYTES>
__device__ __forceinline__ int32_t code_at(const void *v, int64_t rows, int64_t k)
{
    if (k < 0 || k >= rows) return -1;
    if (BYTES == 1) {
        const int8_t c = ((const int8_t *)v)[k];
        return c == (int8_t)OLAP_DIM_NULL_I8 ? -1 : (int32_t)c;
    }
    return ((const int32_t *)v)[k];
}

__device__ __forceinline__ int32_t code_of(const void *v, int64_t rows, int64_t k, int32_t bytes)
{
    return bytes == 1 ? code_at<1>(v, rows, k) : code_at<4>(v, rows, k);
}

template <bool VEC>
__global__ __launch_bounds__(kSnowBS) void k_tri_join(SnowArgs a)
{
    __shared__ unsigned long long lsum[kSnowSlots];
    __shared__ uint32_t lcnt[kSnowSlots];
    for (int g = threadIdx.x; g < a.G; g += kSnowBS) {
        lsum[g] = 0ull;
        lcnt[g] = 0u;
    }
    __syncthreads();
    uint32_t nbad = 0;
    auto add = [&](int32_t g, int32_t price, int32_t disc) {
        if ((uint32_t)g >= (uint32_t)a.G) {
            nbad++;
            return;
        }
        const long long rev = (long long)price * (long long)(100 - disc);
        __hip_atomic_fetch_add(&lsum[g], (unsigned long long)rev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&lcnt[g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    int64_t done = 0;
    if (VEC) {
        const int64_t chunks = a.rows / 4;
        const int64_t stride = (int64_t)gridDim.x * kSnowBS;
        const v4i *sk4 = (const v4i *)a.skey;
        const v4i *ok4 = (const v4i *)a.okey;
        const v4i *pr4 = (const v4i *)a.price;
        const v4i *ds4 = (const v4i *)a.disc;
        for (int64_t c0 = blockIdx.x * (int64_t)kSnowBS + threadIdx.x; c0 < chunks; c0 += stride * kSnowU) {
            v4i sk[kSnowU], ok_[kSnowU];
#pragma unroll
            for (int u = 0; u < kSnowU; u++) {
                const int64_t c = c0 + u * stride;
                sk[u] = c < chunks ? __builtin_nontemporal_load(&sk4[c]) : v4i{-1, -1, -1, -1};
                ok_[u] = c < chunks ? __builtin_nontemporal_load(&ok4[c]) : v4i{-1, -1, -1, -1};
            }
#pragma unroll
            for (int u = 0; u < kSnowU; u++) {
                const int64_t c = c0 + u * stride;
                if (c >= chunks) break;
                int32_t s[4];
                s[0] = code_of(a.svec, a.srows, (int64_t)sk[u].x + a.sbias, a.sbytes);
                s[1] = code_of(a.svec, a.srows, (int64_t)sk[u].y + a.sbias, a.sbytes);
                s[2] = code_of(a.svec, a.srows, (int64_t)sk[u].z + a.sbias, a.sbytes);
                s[3] = code_of(a.svec, a.srows, (int64_t)sk[u].w + a.sbias, a.sbytes);
                // both probes at once: the order keys stream like the supplier
                // keys (a fifth of the rows pass the supplier probe, so almost
                // every line of them is needed), and the order vector is read
                // in key order
                const v4i ok = ok_[u];
                int32_t o[4];
                o[0] = code_of(a.ovec, a.orows, (int64_t)ok.x + a.obias, a.obytes);
                o[1] = code_of(a.ovec, a.orows, (int64_t)ok.y + a.obias, a.obytes);
                o[2] = code_of(a.ovec, a.orows, (int64_t)ok.z + a.obias, a.obytes);
                o[3] = code_of(a.ovec, a.orows, (int64_t)ok.w + a.obias, a.obytes);
                bool m[4];
#pragma unroll
                for (int j = 0; j < 4; j++) m[j] = s[j] != -1 && s[j] == o[j];
                if (!(m[0] || m[1] || m[2] || m[3])) continue;
                const v4i pr = __builtin_nontemporal_load(&pr4[c]);
                const v4i ds = __builtin_nontemporal_load(&ds4[c]);
                if (m[0]) add(s[0], pr.x, ds.x);
                if (m[1]) add(s[1], pr.y, ds.y);
                if (m[2]) add(s[2], pr.z, ds.z);
                if (m[3]) add(s[3], pr.w, ds.w);
            }
        }
        done = chunks * 4;
    }
    for (int64_t i = done + blockIdx.x * (int64_t)kSnowBS + threadIdx.x; i < a.rows;
         i += (int64_t)gridDim.x * kSnowBS) {
        const int32_t s = code_of(a.svec, a.srows, (int64_t)a.skey[i] + a.sbias, a.sbytes);
        if (s == -1) continue;
        const int32_t o = code_of(a.ovec, a.orows, (int64_t)a.okey[i] + a.obias, a.obytes);
        if (o == s) add(s, a.price[i], a.disc[i]);
    }
    for (int o = 32; o > 0; o >>= 1) nbad += __shfl_xor(nbad, o);
    if ((threadIdx.x & 63) == 0 && nbad)
        __hip_atomic_fetch_add(a.bad, nbad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int g = threadIdx.x; g < a.G; g += kSnowBS)
        if (lcnt[g]) {
            __hip_atomic_fetch_add(&a.sum[g], lsum[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&a.cnt[g], (unsigned long long)lcnt[g], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
}

}  // namespace

hipError_t launch_tri_join(const int32_t *skey, int32_t sbias, const void *svec, int64_t srows,
                           int32_t sbytes, const int32_t *okey, int32_t obias, const void *ovec,
                           int64_t orows, int32_t obytes, const int32_t *price, const int32_t *disc,
                           int64_t rows, int32_t G, unsigned long long *sum, unsigned long long *cnt,
                           uint32_t *bad, int num_cus, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    SnowArgs a{skey, okey, price, disc, svec, ovec, srows, orows, sbias, obias, sbytes, obytes,
               rows, G, sum, cnt, bad};
    auto al = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    const bool vec = al(skey) && al(okey) && al(price) && al(disc);
    int64_t g = (rows + (int64_t)kSnowBS * 4 * kSnowU - 1) / ((int64_t)kSnowBS * 4 * kSnowU);
    g = g < 1 ? 1 : g > (int64_t)num_cus * 4 ? (int64_t)num_cus * 4 : g;
    if (vec) k_tri_join<true><<<(int)g, kSnowBS, 0, s>>>(a);
    else k_tri_join<false><<<(int)g, kSnowBS, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace olap
