// olap_snowflake.hip -- TPC-H Q5's fused snowflake probe on gfx950.
//
// Per fact (lineitem) row: the supplier vector gives the supplier's nation
// code for the query's region (NULL otherwise), the orders vector gives the
// ordering customer's nation code for the region and date range; the row
// counts when both are non-NULL and equal (customer and supplier in the same
// nation), and its revenue price x (100 - discount%) goes to that nation's
// int64 slot. Both keys stream through 128-bit loads. The vector with the
// smaller non-NULL fraction (estimated on the device before the probe from
// an evenly strided sample of kSnowSample entries) is
// probed first, and the other only for the rows it passes: in Q5 the orders
// vector passes ~3% of lineitems and is read almost in key order, while the
// supplier vector would be a random gather for every row. Price and
// discount are read only for 4-row chunks with a matching row; the
// per-nation table (up to kSnowSlots slots) sits in LDS.
#include "olap_internal.h"

namespace olap {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kSnowBS = 512;
constexpr int kSnowSlots = 1024;  // largest group count (OLAP_TRI_MAX_GROUPS)
constexpr int kSnowU = 4;  // 128-bit chunks per lane per tile
constexpr int kSnowSample = 65536;  // entries sampled per vector for the probe order

struct SnowArgs {
    const int32_t *skey, *okey, *price, *disc;
    const void *svec, *ovec;
    int64_t srows, orows;
    int32_t sbias, obias, sbytes, obytes;
    int64_t rows;
    int32_t G;
    unsigned long long *sum, *cnt;
    uint32_t *bad;          // rows whose code is outside [0, G)
    const uint32_t *sampled;  // non-NULL entries sampled from the supplier, then the orders vector
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

// non-NULL entries among kSnowSample evenly strided entries of a vector
// (all of them when it is shorter), added to *out
__global__ __launch_bounds__(256) void k_nonnull_sample(const void *v, int64_t rows, int32_t bytes,
                                                        uint32_t *out)
{
    const int64_t n = rows < kSnowSample ? rows : kSnowSample;
    uint32_t c = 0;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        c += code_of(v, rows, n == rows ? i : i * (rows / n), bytes) != -1 ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) __hip_atomic_fetch_add(out, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    // probe the sparser vector first: A is probed for every row, B only
    // where A found a code
    const double fs = a.srows > 0 ? (double)a.sampled[0] / (double)(a.srows < kSnowSample ? a.srows : kSnowSample) : 0.0;
    const double fo = a.orows > 0 ? (double)a.sampled[1] / (double)(a.orows < kSnowSample ? a.orows : kSnowSample) : 0.0;
    const bool ofirst = fo < fs;
    const void *va = ofirst ? a.ovec : a.svec;
    const void *vb = ofirst ? a.svec : a.ovec;
    const int64_t ra = ofirst ? a.orows : a.srows, rb = ofirst ? a.srows : a.orows;
    const int32_t ba = ofirst ? a.obias : a.sbias, bb = ofirst ? a.sbias : a.obias;
    const int32_t ya = ofirst ? a.obytes : a.sbytes, yb = ofirst ? a.sbytes : a.obytes;
    auto probe = [&](int32_t keya, int32_t keyb, int32_t &code) {
        code = code_of(va, ra, (int64_t)keya + ba, ya);
        return code != -1 && code_of(vb, rb, (int64_t)keyb + bb, yb) == code;
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
                // the keys of both steps stream: a few percent of the rows
                // pass the first probe, so most lines of the second step's
                // keys still hold one
                const v4i ka = ofirst ? ok_[u] : sk[u];
                const v4i kb = ofirst ? sk[u] : ok_[u];
                int32_t code[4];
                bool m[4];
#pragma unroll
                for (int j = 0; j < 4; j++) m[j] = probe(ka[j], kb[j], code[j]);
                if (!(m[0] || m[1] || m[2] || m[3])) continue;
                const v4i pr = __builtin_nontemporal_load(&pr4[c]);
                const v4i ds = __builtin_nontemporal_load(&ds4[c]);
                if (m[0]) add(code[0], pr.x, ds.x);
                if (m[1]) add(code[1], pr.y, ds.y);
                if (m[2]) add(code[2], pr.z, ds.z);
                if (m[3]) add(code[3], pr.w, ds.w);
            }
        }
        done = chunks * 4;
    }
    for (int64_t i = done + blockIdx.x * (int64_t)kSnowBS + threadIdx.x; i < a.rows;
         i += (int64_t)gridDim.x * kSnowBS) {
        int32_t code;
        if (probe(ofirst ? a.okey[i] : a.skey[i], ofirst ? a.skey[i] : a.okey[i], code))
            add(code, a.price[i], a.disc[i]);
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
                           uint32_t *bad, uint32_t *sampled, int num_cus, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    // the non-NULL fraction of each vector, for the probe order
    hipError_t e = launch_zero_u32(sampled, 2, s);
    if (e != hipSuccess) return e;
    if (srows > 0) k_nonnull_sample<<<kSnowSample / 256 / 4, 256, 0, s>>>(svec, srows, sbytes, sampled);
    if (orows > 0) k_nonnull_sample<<<kSnowSample / 256 / 4, 256, 0, s>>>(ovec, orows, obytes, sampled + 1);
    SnowArgs a{skey, okey, price, disc, svec, ovec, srows, orows, sbias, obias, sbytes, obytes,
               rows, G, sum, cnt, bad, sampled};
    auto al = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    const bool vec = al(skey) && al(okey) && al(price) && al(disc);
    int64_t g = (rows + (int64_t)kSnowBS * 4 * kSnowU - 1) / ((int64_t)kSnowBS * 4 * kSnowU);
    g = g < 1 ? 1 : g > (int64_t)num_cus * 4 ? (int64_t)num_cus * 4 : g;
    if (vec) k_tri_join<true><<<(int)g, kSnowBS, 0, s>>>(a);
    else k_tri_join<false><<<(int)g, kSnowBS, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace olap
