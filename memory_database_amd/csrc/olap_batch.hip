// olap_batch.hip -- the batched form: up to 64 star-join group-by queries
// evaluated in one persistent scan of the shared fact table.
//
// Per dimension table t the batch keeps a 64-bit query mask per dimension
// row (bit q: query q passes row r of t, or does not probe t). Per fact row
// the kernel gathers one mask and one group attribute per probed table --
// from LDS for the small tables, from L2/HBM for the large ones -- ANDs the
// masks, and for every surviving query composes its gid from the attributes
// (code = attr mod group_mod by an exact multiply-high fast modulo) and adds
// the measure into the query's group table: in LDS when the autotuned LDS
// budget holds it, else with global atomics on the query's output slice.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "olap_internal.h"

namespace olap {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr uint32_t kNoLds = 0xffffffffu;

// one query as the batched kernels read it (128 bytes). The group-by terms
// are kept by table (the kernels' numbering): gid = sum over the four tables
// of (attr mod g[t]) * fac[t], a table the query does not group by, or
// groups by modulo 1, having factor 0 (it adds 0 whatever its attribute)
struct BQDesc {
    uint32_t nsteps;    // probe steps (the kernels compute all four table terms)
    uint32_t lds_tab;   // word offset of its LDS group table, kNoLds when in HBM
    uint32_t G;
    uint32_t out_off;   // first slot in the concatenated outputs
    uint32_t fac[4];    // words 4..7
    uint32_t g[4];      // words 8..11: group_mod (0 or 1: code 0)
    uint32_t m32[4];    // words 12..15: 32-bit modulo multiplier (attributes below 2^31)
    uint64_t M[4];      // words 16..23: fast-modulo multiplier UINT64_MAX / g + 1 (0 for g 0)
    uint32_t pbase;     // word 24: first staged-aggregation partition of the query
    uint32_t pbase1;    // the same for the one-pass records of the batch's format
    uint32_t seg_base;  // segmented kernel: first position of the query's event segment
    uint32_t seg_cap;   // and its capacity (a multiple of 64, at least 64)
    uint32_t sh4;       // word 28: the shifts of the 32-bit modulo, 8 bits a table
    uint32_t pad[3];
};
static_assert(sizeof(BQDesc) == 128, "descriptor layout");
constexpr uint32_t kDescWords = sizeof(BQDesc) / 4;

struct BatchArgs {
    const int32_t *fk[4];
    const int32_t *m1;
    const int32_t *m2;
    uint64_t rows;
    uint64_t ntiles;
    const unsigned long long *gmask[4];
    const uint4 *grec[4];   // per dimension row: mask (2 words), attribute, 0
    const uint4 *grec_any;  // the records of some probed table (a valid address)
    uint32_t dim_rows[4];
    uint32_t mask_lds[4];   // word offset of the LDS copy or kNoLds
    uint32_t attr_lds[4];
    unsigned long long nouse[4];
    uint32_t used;          // bit t: some query probes table t
    uint32_t nq;
    const BQDesc *desc;
    uint32_t desc_lds;      // word offset of the descriptors in LDS
    uint32_t tab_begin;     // LDS words [tab_begin, tab_end) hold query tables
    uint32_t tab_end;
    unsigned long long *gsum;
    unsigned long long *gcnt;
    uint32_t *err;
    StageArgs stage;
    // event-grouped kernel: LDS word offsets of the per-query counters
    // (then cursors, starts and the total), the staged attributes [R][4],
    // the staged measures [R] (m1, then m2) and the event list of ev_cap
    uint32_t cnt_lds;
    uint32_t row_lds;
    uint32_t m_lds;
    uint32_t ev_lds;
    uint32_t ev_cap;
    // one-pass staging: per-partition chunk state in LDS (u32 current and
    // spare chunk ids, then u32 fill, nparts each)
    ChunkArgs chunk;
    uint32_t pc_lds;
    uint64_t qvalid;        // bit q: q < nq
    // every attribute below 2^31 and every G below 2^24: the group-by terms
    // take the 32-bit multiply-high modulo (fast_code)
    uint32_t fastmod;
};

struct MaskPred {
    int32_t q;
    int32_t lo;
    int32_t hi;
    int32_t pad;
    const int8_t *pass;  // a general predicate's pass vector (NULL 127 fails), or null
};

// explicit address spaces keep the LDS and the HBM atomics apart (a generic
// 64-bit atomic would need a run-time address-space check)
typedef __attribute__((address_space(1))) unsigned long long g_u64;
typedef __attribute__((address_space(3))) unsigned long long l_u64;
typedef __attribute__((address_space(3))) uint32_t l_u32;
// descriptors read with a wave-uniform query index go through the constant
// address space, so they become scalar loads
typedef __attribute__((address_space(4))) const struct BQDesc c_desc;

__device__ __forceinline__ void b_atomic_dev(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_fetch_add((g_u64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void b_atomic_lds(uint32_t *tab, uint32_t G, uint32_t gid,
                                             unsigned long long v)
{
    l_u64 *ls = (l_u64 *)tab;
    l_u32 *lc = (l_u32 *)(tab + 2u * G);
    __hip_atomic_fetch_add(ls + gid, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(lc + gid, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------- masks

// (neg: set when an attribute is negative, i.e. 2^31 or more as the kernels
// read it). The pass counts are summed per workgroup in LDS (one wave
// popcount per predicate and wave, no contention between workgroups) and
// added to `pass` once per workgroup; the grid is a few workgroups per CU
// striding over the rows.
constexpr int kMaskThreads = 256;
__global__ __launch_bounds__(kMaskThreads) void k_batch_masks(const int32_t *filter, const int32_t *attr, uint32_t D,
                                                              const MaskPred *preds, int np, unsigned long long base,
                                                              unsigned long long *mask, uint4 *rec, uint32_t *pass,
                                                              uint32_t *neg)
{
    __shared__ uint32_t cnt[4 * OLAP_BATCH_MAX];
    for (int i = threadIdx.x; i < np; i += kMaskThreads) cnt[i] = 0u;
    __syncthreads();
    bool any_neg = false;
    for (uint32_t r0 = blockIdx.x * kMaskThreads; r0 < D; r0 += gridDim.x * kMaskThreads) {
        const uint32_t r = r0 + threadIdx.x;
        const bool in = r < D;
        const int32_t f = in ? filter[r] : 0;
        unsigned long long m = base;
        for (int i = 0; i < np; i++) {
            const int8_t *pv = preds[i].pass;
            const bool ok = in && f >= preds[i].lo && f < preds[i].hi &&
                            (!pv || pv[r] != (int8_t)OLAP_DIM_NULL_I8);
            if (ok) m |= 1ull << preds[i].q;
            const unsigned long long b = __ballot(ok);
            if ((threadIdx.x & 63) == 0 && b) atomicAdd(&cnt[i], (uint32_t)__popcll(b));
        }
        if (in) {
            const int32_t at = attr ? attr[r] : 0;
            mask[r] = m;
            rec[r] = make_uint4((uint32_t)m, (uint32_t)(m >> 32), (uint32_t)at, 0u);
            any_neg |= at < 0;
        }
    }
    if (__ballot(any_neg) && (threadIdx.x & 63) == 0) atomicOr(neg, 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += kMaskThreads)
        if (cnt[i]) atomicAdd(&pass[i], cnt[i]);
}

// min and max of a measure column: a grid-stride pass with 16-byte loads
// (the column is 16-byte aligned), a wave reduction, one atomic pair a wave
__global__ void k_minmax_init(int32_t *out)
{
    out[0] = INT32_MAX;
    out[1] = INT32_MIN;
}

__global__ void k_minmax(const int32_t *m, uint64_t n, int32_t *out)
{
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    const uint64_t n4 = n / 4u;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const v4i v = __builtin_nontemporal_load((const v4i *)m + i);
        lo = min(min(lo, v.x), min(v.y, min(v.z, v.w)));
        hi = max(max(hi, v.x), max(v.y, max(v.z, v.w)));
    }
    for (uint64_t i = 4u * n4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        lo = min(lo, m[i]);
        hi = max(hi, m[i]);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d));
        hi = max(hi, __shfl_xor(hi, d));
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicMin(&out[0], lo);
        atomicMax(&out[1], hi);
    }
}

// the range of n > 0 values of m into out (synchronises s)
hipError_t measure_range(const int32_t *m, uint64_t n, int32_t *dev2, int32_t out[2], hipStream_t s)
{
    k_minmax_init<<<1, 1, 0, s>>>(dev2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n / 4u + 255u) / 256u, 4096u));
    k_minmax<<<grid, 256, 0, s>>>(m, n, dev2);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, dev2, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

// ---------------------------------------------------------------- scan

template <bool VEC>
__device__ __forceinline__ v4i bload4(const int32_t *__restrict__ p, uint64_t r, uint64_t rows,
                                      uint32_t &present)
{
    v4i v = {0, 0, 0, 0};
    if (r + 4 <= rows) {
        present = 0xFu;
        v = __builtin_nontemporal_load((const v4i *)(p + r));
    } else {
        present = 0u;
        if (r < rows) { v.x = p[r]; present |= 1u; }
        if (r + 1 < rows) { v.y = p[r + 1]; present |= 2u; }
        if (r + 2 < rows) { v.z = p[r + 2]; present |= 4u; }
    }
    return v;
}

// MODE AGG_DIRECT: queries without an LDS table use global atomics.
// AGG_COUNT / AGG_SCATTER: the two passes of the staged aggregation; queries
// with an LDS table are aggregated in the COUNT pass only.
template <int BS, int U, int MODE>
__global__ __launch_bounds__(BS) void k_batch(BatchArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];
    stage_begin<MODE>(a.stage, lds);
    // stage small tables' masks/attributes and the descriptors; zero tables
    for (int t = 0; t < 4; t++) {
        if (a.mask_lds[t] != kNoLds) {
            const uint32_t *src = (const uint32_t *)a.gmask[t];
            for (uint32_t i = threadIdx.x; i < 2u * a.dim_rows[t]; i += BS) lds[a.mask_lds[t] + i] = src[i];
        }
        if (a.attr_lds[t] != kNoLds)  // the attributes as the records hold them
            for (uint32_t i = threadIdx.x; i < a.dim_rows[t]; i += BS) lds[a.attr_lds[t] + i] = a.grec[t][i].z;
    }
    {
        const uint32_t *src = (const uint32_t *)a.desc;
        const uint32_t n = a.nq * kDescWords;
        for (uint32_t i = threadIdx.x; i < n; i += BS) lds[a.desc_lds + i] = src[i];
    }
    for (uint32_t i = a.tab_begin + threadIdx.x; i < a.tab_end; i += BS) lds[i] = 0u;
    __syncthreads();
    // the uniform flush loop reads descriptors from the global copy
    const BQDesc *D = a.desc;
    const uint64_t rows = a.rows;
    const uint32_t tid = threadIdx.x;
    constexpr uint64_t TILE = (uint64_t)BS * 4u * U;
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile * TILE;
        v4i f[U][4];
        uint32_t present[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t r = base + (uint64_t)(u * BS + tid) * 4u;
            present[u] = 0;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                f[u][t] = v4i{0, 0, 0, 0};
                if ((a.used >> t) & 1u) f[u][t] = bload4<true>(a.fk[t], r, rows, present[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t r = base + (uint64_t)(u * BS + tid) * 4u;
            unsigned long long alive[4];
            uint32_t at[4][4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                alive[k] = ((present[u] >> k) & 1u) ? ~0ull : 0ull;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    at[t][k] = 0u;
                    if (((a.used >> t) & 1u) && alive[k]) {
                        const uint32_t key = (uint32_t)f[u][t][k];
                        if (key < a.dim_rows[t]) {
                            if (a.mask_lds[t] != kNoLds) {
                                const l_u64 *w = (const l_u64 *)((const l_u32 *)lds + a.mask_lds[t]);
                                alive[k] &= w[key];
                                at[t][k] = ((const l_u32 *)lds)[a.attr_lds[t] + key];
                            } else {
                                // one 16-byte gather: mask (2 words), attribute
                                const uint4 rec = a.grec[t][key];
                                alive[k] &= (unsigned long long)rec.x | ((unsigned long long)rec.y << 32);
                                at[t][k] = rec.z;
                            }
                        } else {
                            alive[k] &= a.nouse[t];
                        }
                    }
                }
            }
            if ((alive[0] | alive[1] | alive[2] | alive[3]) == 0ull) continue;
            uint32_t dummy;
            const v4i m1 = bload4<true>(a.m1, r, rows, dummy);
            v4i m2 = {0, 0, 0, 0};
            if (a.m2) m2 = bload4<true>(a.m2, r, rows, dummy);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                unsigned long long al = alive[k];
                const unsigned long long mv =
                    (unsigned long long)((long long)m1[k] + (a.m2 ? (long long)m2[k] : 0ll));
                while (al) {
                    const int q = __ffsll(al) - 1;
                    al &= al - 1ull;
                    // descriptor words of query q (BQDesc layout) read from the
                    // LDS copy through an explicit address-space-3 pointer
                    const l_u32 *dw = (const l_u32 *)lds + a.desc_lds + kDescWords * (uint32_t)q;
                    const uint32_t lds_tab = dw[1], G = dw[2];
                    uint32_t gid = 0;
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const uint64_t M = ((const l_u64 *)(dw + 16))[t];
                        const uint32_t code = (uint32_t)__umul64hi(M * (uint64_t)at[t][k], (uint64_t)dw[8 + t]);
                        gid += code * dw[4 + t];
                    }
                    if (gid >= G) {
                        __hip_atomic_fetch_add(a.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        continue;
                    }
                    if (lds_tab != kNoLds) {
                        if (MODE != AGG_SCATTER) b_atomic_lds(lds + lds_tab, G, gid, mv);
                    } else if (MODE == AGG_DIRECT) {
                        b_atomic_dev(&a.gsum[dw[3] + gid], mv);
                        b_atomic_dev(&a.gcnt[dw[3] + gid], 1ull);
                    } else if (MODE == AGG_COUNT) {
                        stage_count(a.stage, lds, dw[24] + (gid >> kPartBits));
                    } else {
                        stage_scatter(a.stage, lds, dw[24] + (gid >> kPartBits),
                                      gid & (kPartSlots - 1u), m1[k], a.m2 ? m2[k] : 0);
                    }
                }
            }
        }
    }
    __syncthreads();
    if (MODE == AGG_COUNT) stage_end<MODE>(a.stage, lds);
    if (MODE == AGG_SCATTER) return;
    for (uint32_t q = 0; q < a.nq; q++) {
        const BQDesc &d = D[q];
        if (d.lds_tab == kNoLds) continue;
        const l_u64 *ls = (const l_u64 *)((l_u32 *)lds + d.lds_tab);
        const l_u32 *lc = (l_u32 *)lds + d.lds_tab + 2u * d.G;
        for (uint32_t g = tid; g < d.G; g += BS) {
            const uint32_t c = lc[g];
            if (c) {
                b_atomic_dev(&a.gcnt[d.out_off + g], (unsigned long long)c);
                b_atomic_dev(&a.gsum[d.out_off + g], ls[g]);
            }
        }
    }
}

// ---------------------------------------------------------- grouped events
// Event-grouped form of the scan. The per-lane loop above leaves most lanes
// idle (a wave waits for its busiest lane) and every lane reads a different
// query's descriptor. Here a tile of R = 2 * BS rows goes through five steps:
//  1. rows: each lane loads the foreign keys of its two rows, gathers the
//     per-table query masks and attributes, and stages the attributes and
//     measures of its rows in LDS;
//  2. count: one LDS atomic per (row, query) match on the query's counter;
//  3. offsets: exclusive sums over the queries;
//  4. list: a cursor atomic per match places it in an LDS event list grouped
//     by query;
//  5. events: every query's segment is padded to whole waves, so each wave
//     takes 64 events of one query: its descriptor is uniform, and a query
//     of one aggregation partition takes one counter/cursor atomic per wave.
// A tile with more matches than the list holds runs steps 4-5 in rounds, each
// over a run of whole query segments. In the one-pass staged mode every event
// of step 5 writes its record, and a sixth step retires the full record
// chunks: between rounds, and after a tile's last round while the next
// tile's offsets are computed.

__device__ __forceinline__ uint32_t lane_rank(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Workgroup barrier for LDS hand-offs: waits for this wave's LDS operations
// (lgkmcnt 0), not for its global stores, which no thread of the workgroup
// reads back inside the kernel (records, chunk tables and atomics are read by
// later kernels). The signal fences keep the compiler from moving memory
// operations across it.
__device__ __forceinline__ void lds_barrier()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt 63, expcnt 7, lgkmcnt 0
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// w rotated right by r (0..63)
__device__ __forceinline__ unsigned long long rotr64(unsigned long long w, uint32_t r)
{
    return (w >> r) | (w << ((64u - r) & 63u));
}

// w rotated left by r (0..63)
__device__ __forceinline__ unsigned long long rotl64(unsigned long long w, uint32_t r)
{
    return (w << r) | (w >> ((64u - r) & 63u));
}

typedef int v2i_a4 __attribute__((ext_vector_type(2), aligned(4)));  // a pair at any 4-byte offset

// (attr mod g) * fac for attr < 2^31, g < 2^24 and fac < 2^24 (m and s from
// fastmod_magic; g 0 or 1 has m 0 and fac 0): the quotient is
// mulhi(attr, m) >> s. For g not a power of two, s = floor(log2 g) and
// m = ceil(2^(32+s) / g) = (2^(32+s) + e) / g with 0 < e < g; for
// attr = q g + r, attr m / 2^(32+s) = q + (r + attr e / 2^(32+s)) / g, and
// attr e / 2^(32+s) < attr g / 2^(32+s) < attr / 2^31 < 1, so the floor is
// q. For g = 2^s, m = 2^31 and the shift is s - 1.
__device__ __forceinline__ uint32_t fast_term(uint32_t attr, uint32_t m, uint32_t s, uint32_t g, uint32_t fac)
{
    const uint32_t q = __umulhi(attr, m) >> s;
    return __umul24(attr - q * g, fac);
}

// The aggregation of one event whose gid is known (wave-uniform G and
// lds_tab; pbase the query's first partition for MODE): its LDS table,
// global atomics, or a staged record (the COUNT/SCATTER passes or a one-pass
// chunk). `valid` false is a padding lane: it takes part in the wave's
// ballots only.
template <int MODE, bool CPT = false>
__device__ __forceinline__ void event_aggregate(const BatchArgs &a, bool valid, uint32_t gid, uint32_t G,
                                                uint32_t lds_tab, uint32_t pbase, uint32_t out_off, uint32_t m1v,
                                                uint32_t m2v, uint32_t *lds, stage_l_u32 *pc_cur,
                                                stage_l_u32 *pc_spare, stage_l_u32 *pc_fill, stage_l_u32 *scur,
                                                const stage_l_u64 *sstart, uint32_t CH, uint32_t lane)
{
    const bool ok = valid && gid < G;
    if (valid && gid >= G)
        __hip_atomic_fetch_add(a.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long mv = (long long)(int32_t)m1v + (CPT ? 0ll : (long long)(int32_t)m2v);
    const bool compact = MODE == AGG_CHUNK && (CPT || a.chunk.fmt == REC_COMPACT);
    const uint32_t pbits = compact ? kCompactPartBits : kPartBits;
    const uint32_t part = pbase + (gid >> pbits);
    const uint32_t glow = gid & ((1u << pbits) - 1u);
    if (lds_tab != kNoLds) {
        if (MODE != AGG_SCATTER && ok) b_atomic_lds(lds + lds_tab, G, gid, (unsigned long long)mv);
    } else if (MODE == AGG_DIRECT) {
        if (ok) {
            b_atomic_dev(&a.gsum[out_off + gid], (unsigned long long)mv);
            b_atomic_dev(&a.gcnt[out_off + gid], 1ull);
        }
    } else {
        // (4-byte records hold one measure: m1 - mbase in 32 bits)
        const uint32_t moff = m1v - (uint32_t)a.chunk.mbase;
        // a measure outside the range 4-byte records were packed for (the
        // columns changed since the range was taken) is added directly and
        // takes no record slot
        bool rec = ok;
        if (compact && ok && ((int32_t)m1v < a.chunk.mbase || moff >= (1u << kCompactMeasureBits))) {
            const unsigned long long o = a.chunk.part_out[part] + glow;
            b_atomic_dev(&a.gsum[o], (unsigned long long)mv);
            b_atomic_dev(&a.gcnt[o], 1ull);
            __hip_atomic_fetch_add(a.chunk.overflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            rec = false;
        }
        // counters: the COUNT pass's per-partition counts, the SCATTER
        // pass's cursors, or the one-pass chunk fills
        stage_l_u32 *ctr = MODE == AGG_CHUNK ? pc_fill : scur;
        // the partition's chunk ids (fixed during the events step), read
        // before the fill atomic
        uint32_t id_cur = 0u, id_spare = 0u;
        if (MODE == AGG_CHUNK && rec) {
            id_cur = pc_cur[part];
            id_spare = pc_spare[part];
        }
        uint32_t pos = 0u;
        if (G <= (1u << pbits)) {
            // one partition: one counter / cursor atomic for the wave
            const unsigned long long act = __ballot(rec);
            if (act) {
                const int leader = __ffsll(act) - 1;
                // the leader's rank is 0: adding it keeps the address per
                // lane, so the compiler does not rewrite this single-lane
                // atomic as a wave reduction
                const uint32_t rank = lane_rank(act);
                uint32_t at = 0u;
                if ((int)lane == leader)
                    at = __hip_atomic_fetch_add(ctr + pbase + rank, (uint32_t)__popcll(act),
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                pos = (uint32_t)__builtin_amdgcn_readlane((int)at, leader) + rank;
            }
        } else if (rec) {
            pos = __hip_atomic_fetch_add(ctr + part, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (MODE == AGG_CHUNK && rec) {
            // a round adds at most R = CH records to a partition whose fill
            // is below CH: pos < 2 CH
            const uint32_t id = pos < CH ? id_cur : id_spare;
            const uint32_t off = pos < CH ? pos : pos - CH;
            const unsigned long long base = (unsigned long long)id << a.chunk.ch_bits;
            if (id == kNoChunk) {
                // no chunk (the pool is used up): added directly
                const unsigned long long o = a.chunk.part_out[part] + glow;
                b_atomic_dev(&a.gsum[o], (unsigned long long)mv);
                b_atomic_dev(&a.gcnt[o], 1ull);
                __hip_atomic_fetch_add(a.chunk.overflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if (compact) {
                ((uint32_t *)a.chunk.pool)[base + off] = (glow << kCompactMeasureBits) | moff;
            } else if (CPT) {
                // (not reached: CPT kernels run COMPACT pools only)
            } else if (a.m2) {
                ((uint4 *)a.chunk.pool)[base + off] = make_uint4(glow, m1v, m2v, 0u);
            } else {
                // one measure: a chunk holds CH u32 measures, then CH u16
                // slots (6 bytes a record)
                char *cp = (char *)a.chunk.pool + 6ull * base;
                ((uint32_t *)cp)[off] = m1v;
                ((uint16_t *)(cp + 4u * CH))[off] = (uint16_t)glow;
            }
        }
        if (MODE == AGG_SCATTER && ok) {
            const unsigned long long idx = sstart[part] + pos;
            if (idx >= a.stage.cap) {
                stage_overflow(a.stage, part, gid & (kPartSlots - 1u), mv);
            } else if (a.stage.rec4) {
                a.stage.rec4[idx] = make_uint4(gid & (kPartSlots - 1u), m1v, m2v, 0u);
            } else {
                a.stage.rec2[idx] = make_uint2(gid & (kPartSlots - 1u), m1v);
            }
        }
    }
}

// One event of a wave-uniform query q (descriptor d): the gid from the
// row's staged attributes and the query's group-by terms, then its
// aggregation. FAST: 1 the 32-bit modulo, 0 the 64-bit one, -1 as a.fastmod
// says. The descriptor words every path needs are read before any branch,
// so that their scalar loads are issued, and waited for, together.
template <int MODE, int FAST = -1, bool CPT = false>
__device__ __forceinline__ void batch_event(const BatchArgs &a, const c_desc &d, bool valid, uint32_t row,
                                            uint32_t *lds, const l_u32 *rattr, const l_u32 *rm, uint32_t R,
                                            stage_l_u32 *pc_cur, stage_l_u32 *pc_spare, stage_l_u32 *pc_fill,
                                            stage_l_u32 *scur, const stage_l_u64 *sstart, uint32_t CH,
                                            uint32_t lane)
{
    // the row's attributes of the four tables and its measures (at fixed
    // LDS offsets, issued with the descriptor's scalar loads), then the
    // group-by terms of every table (a table the query does not group by
    // adds 0, whatever the staged word holds)
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) const v4u l_cu4;
    const v4u a4 = ((l_cu4 *)rattr)[row];
    const uint32_t av[4] = {a4.x, a4.y, a4.z, a4.w};
    const uint32_t m1v = rm[row];
    const uint32_t m2v = !CPT && a.m2 ? rm[R + row] : 0u;
    const uint32_t G = d.G;
    const uint32_t lds_tab = d.lds_tab;
    const uint32_t pbase = MODE == AGG_CHUNK ? d.pbase1 : d.pbase;
    uint32_t gid = 0u;
    if (FAST < 0 ? a.fastmod != 0u : FAST == 1) {
#pragma unroll
        for (int t = 0; t < 4; t++) gid += fast_term(av[t], d.m32[t], (d.sh4 >> (8 * t)) & 31u, d.g[t], d.fac[t]);
    } else {
#pragma unroll
        for (int t = 0; t < 4; t++)
            gid += (uint32_t)__umul64hi(d.M[t] * (uint64_t)av[t], (uint64_t)d.g[t]) * d.fac[t];
    }
    event_aggregate<MODE, CPT>(a, valid, gid, G, lds_tab, pbase, d.out_off, m1v, m2v, lds, pc_cur, pc_spare, pc_fill,
                          scur, sstart, CH, lane);
}

template <int BS, int MODE>
__global__ __launch_bounds__(BS) void k_batch_ev(BatchArgs a)
{
    constexpr uint32_t R = 2u * BS;
    extern __shared__ __align__(16) uint32_t lds[];
    stage_begin<MODE>(a.stage, lds);
    for (int t = 0; t < 4; t++) {
        if (a.mask_lds[t] != kNoLds) {
            const uint32_t *src = (const uint32_t *)a.gmask[t];
            for (uint32_t i = threadIdx.x; i < 2u * a.dim_rows[t]; i += BS) lds[a.mask_lds[t] + i] = src[i];
        }
        if (a.attr_lds[t] != kNoLds)  // the attributes as the records hold them
            for (uint32_t i = threadIdx.x; i < a.dim_rows[t]; i += BS) lds[a.attr_lds[t] + i] = a.grec[t][i].z;
    }
    // (descriptors are read from global memory with scalar loads: no LDS copy)
    for (uint32_t i = a.tab_begin + threadIdx.x; i < a.tab_end; i += BS) lds[i] = 0u;
    if (threadIdx.x < 64u) lds[a.cnt_lds + threadIdx.x] = 0u;
    // one-pass staging: per partition a current chunk and a spare (chunk
    // ids, kNoChunk when the pool ran out) and the records handed out in the
    // current chunk (positions past its end go to the spare); a COMPACT pool
    // has partitions of 2^12 slots
    const uint32_t np = a.chunk.nparts;
    const uint32_t CH = 1u << a.chunk.ch_bits;
    stage_l_u32 *pc_cur = (stage_l_u32 *)lds + a.pc_lds;
    stage_l_u32 *pc_spare = pc_cur + np;
    stage_l_u32 *pc_fill = pc_spare + np;
    if (MODE == AGG_CHUNK) {
        for (uint32_t p = threadIdx.x; p < np; p += BS) {
            pc_cur[p] = chunk_claim(a.chunk, p);
            pc_spare[p] = chunk_claim(a.chunk, p);
            pc_fill[p] = 0u;
        }
    }
    __syncthreads();
    l_u32 *L = (l_u32 *)lds;
    l_u32 *qcnt = L + a.cnt_lds;          // [64] matches per query in the tile
    l_u32 *qcur = qcnt + 64u;             // [64] list cursors per query
    l_u32 *qreal = qcur + 64u;            // [64] matches per query (kept for the padding)
    l_u32 *seg = qreal + 64u;             // [64] padded query starts, [64] the padded total
    l_u32 *rattr = L + a.row_lds;         // [R] the row's four attributes (16 bytes a row)
    l_u32 *rm = L + a.m_lds;              // [R] m1, then [R] m2
    l_u32 *ev = L + a.ev_lds;             // [ev_cap]
    stage_l_u32 *scur = (stage_l_u32 *)lds + a.stage.cur_lds;
    const stage_l_u64 *sstart = (const stage_l_u64 *)((stage_l_u32 *)lds + a.stage.start_lds);
    const BQDesc *D = a.desc;
    const uint64_t rows = a.rows;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    // the columns of a tile are loaded one tile ahead: the next tile's keys
    // and measures are issued after this tile's rows step and arrive while
    // its counts, lists and events run (the barriers wait for LDS only). The
    // one-pass form, whose LDS already limits it to two workgroups per CU,
    // also gathers the next tile's records of dimension rows outside LDS
    // before this tile's events (32 more registers per lane, which would cost
    // the other forms occupancy)
    //
    // The loads are unconditional, so the loaded registers need no copies
    // (a copy would wait for the load at once): a pair that reaches past the
    // end reads the last pair in bounds, rows - 2 .. rows - 1 (rows >= 2; a
    // 4-byte-aligned load), and takes its second element for its first row;
    // rows past the end are masked out by al0/al1. The last tile loads
    // itself again and unused columns read m1.
    //
    // A gather takes the vector memory path one cache line per lane, and a
    // wave stalls while that path is full: gathers issued by every wave at
    // once cost each of them the whole burst. So the gathers (table t, row
    // j of the lane's pair) are spread over three points of the tile:
    //   A, after the offsets step (the keys have had the count step to
    //      arrive): (0, 0), (1, 0);
    //   B, before the barrier that ends the first round's list step:
    //      (2, 0), (3, 0), (0, 1), (1, 1);
    //   C, before the barrier that ends its events step: (2, 1), (3, 1).
    // At B and C a wave that issues them would otherwise wait for slower
    // waves. The host numbers the tables outside LDS 0, 2, 1, 3 in turn:
    // 2-4 such tables put 1/2/1, 2/3/1 or 2/4/2 gathers off HBM at A/B/C
    // (one: 1/1/0); a gather off a table kept in LDS reads record 0 of the
    // table grec_any (one cache line for the wave).
    constexpr bool kGatherAhead = MODE == AGG_CHUNK;
    const uint64_t rlast = rows - 2u;
    const int32_t *colp[6];
#pragma unroll
    for (int t = 0; t < 4; t++) colp[t] = (a.used >> t) & 1u ? a.fk[t] : a.m1;
    colp[4] = a.m1;
    colp[5] = a.m2 ? a.m2 : a.m1;
    int2 nk[4], nm1, nm2;
    uint4 ng[4][2];
    auto load_tile = [&](uint64_t tl) {
        const uint64_t r = tl * R + 2u * tid;
        const uint64_t rl = min(r, rlast);
        const bool shifted = r != rl;
        auto ld2 = [&](const int32_t *p) {
            const v2i_a4 v = __builtin_nontemporal_load((const v2i_a4 *)(p + rl));
            return make_int2(shifted ? v.y : v.x, v.y);
        };
#pragma unroll
        for (int t = 0; t < 4; t++) nk[t] = ld2(colp[t]);
        nm1 = ld2(colp[4]);
        nm2 = ld2(colp[5]);
    };
    auto gather = [&](int t, int j) {
        if (!kGatherAhead) return;
        const bool hbm = ((a.used >> t) & 1u) && a.mask_lds[t] == kNoLds;
        const uint4 *g = hbm ? a.grec[t] : a.grec_any;
        const uint32_t kmax = hbm ? max(a.dim_rows[t], 1u) - 1u : 0u;
        ng[t][j] = g[min((uint32_t)(j ? nk[t].y : nk[t].x), kmax)];
    };
    auto gather_a = [&]() { gather(0, 0); gather(1, 0); };
    auto gather_b = [&]() { gather(2, 0); gather(3, 0); gather(0, 1); gather(1, 1); };
    auto gather_c = [&]() { gather(2, 1); gather(3, 1); };
    if (blockIdx.x < a.ntiles) {
        load_tile(blockIdx.x);
        gather_a();
        gather_b();
        gather_c();
    }
    // 6. (one-pass staging) full chunks retire (fill CH), the spare becomes
    // the current chunk and a new spare is claimed. It runs between the
    // rounds of a tile, and for the last round of a tile while the next
    // tile's offsets are computed (no step in between touches the chunk
    // state; the barrier after the offsets orders it before the events)
    auto retire_chunks = [&]() {
        for (uint32_t p = tid; p < np; p += BS) {
            const uint32_t f = pc_fill[p];
            if (f >= CH) {
                const uint32_t cur = pc_cur[p];
                if (cur != kNoChunk) a.chunk.chunk_fill[cur] = CH;
                pc_cur[p] = pc_spare[p];
                pc_fill[p] = f - CH;
                pc_spare[p] = chunk_claim(a.chunk, p);
            }
        }
    };
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t r0 = tile * R + 2u * tid;
        const uint64_t next = tile + gridDim.x < a.ntiles ? tile + gridDim.x : tile;
        int2 ck[4];
        uint4 cg[4][2];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            ck[t] = nk[t];
            cg[t][0] = ng[t][0];
            cg[t][1] = ng[t][1];
        }
        const int2 cm1 = nm1, cm2 = nm2;
        // 1. rows
        unsigned long long al0 = r0 < rows ? ~0ull : 0ull;
        unsigned long long al1 = r0 + 1 < rows ? ~0ull : 0ull;
        uint32_t atr[2][4] = {{0u, 0u, 0u, 0u}, {0u, 0u, 0u, 0u}};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            uint32_t at0 = 0u, at1 = 0u;
            if ((a.used >> t) & 1u) {
                const int2 k = ck[t];
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const uint32_t key = (uint32_t)(j ? k.y : k.x);
                    unsigned long long &al = j ? al1 : al0;
                    uint32_t &at = j ? at1 : at0;
                    if (al) {
                        if (key < a.dim_rows[t]) {
                            if (a.mask_lds[t] != kNoLds) {
                                const l_u64 *w = (const l_u64 *)(L + a.mask_lds[t]);
                                al &= w[key];
                                at = L[a.attr_lds[t] + key];
                            } else {
                                const uint4 rec = kGatherAhead ? cg[t][j] : a.grec[t][key];
                                al &= (unsigned long long)rec.x | ((unsigned long long)rec.y << 32);
                                at = rec.z;
                            }
                        } else {
                            al &= a.nouse[t];
                        }
                    }
                }
                atr[0][t] = at0;
                atr[1][t] = at1;
            }
        }
        {
            // one 16-byte word a row: the events step reads a row's four
            // attributes with one LDS read
            typedef unsigned int v4u __attribute__((ext_vector_type(4)));
            typedef __attribute__((address_space(3))) v4u l_u4;
            l_u4 *r4 = (l_u4 *)rattr;
            r4[2u * tid] = v4u{atr[0][0], atr[0][1], atr[0][2], atr[0][3]};
            r4[2u * tid + 1u] = v4u{atr[1][0], atr[1][1], atr[1][2], atr[1][3]};
        }
        if (al0 | al1) {
            rm[2u * tid] = (uint32_t)cm1.x;
            rm[2u * tid + 1u] = (uint32_t)cm1.y;
            if (a.m2) {
                rm[R + 2u * tid] = (uint32_t)cm2.x;
                rm[R + 2u * tid + 1u] = (uint32_t)cm2.y;
            }
        }
        load_tile(next);
        // 2. count: one LDS atomic per match on its query's counter; each lane
        // walks its bits from its own lane number on, so the lanes of one
        // atomic instruction mostly hit different counters
        for (int j = 0; j < 2; j++) {
            // the mask rotated right by the lane number: bit b is query
            // (b + lane) mod 64, and w & (w - 1) drops the bit just taken
            for (unsigned long long w = rotr64(j ? al1 : al0, lane); w; w &= w - 1ull) {
                const uint32_t q = ((uint32_t)__builtin_ctzll(w) + lane) & 63u;
                __hip_atomic_fetch_add(qcnt + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        lds_barrier();
        // 3. offsets: every query's segment is padded to whole waves
        if (tid < 64u) {
            const uint32_t c = qcnt[tid];
            const uint32_t padded = (c + 63u) & ~63u;
            uint32_t incl = padded;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t v = (uint32_t)__shfl_up((int)incl, d);
                if (tid >= (uint32_t)d) incl += v;
            }
            seg[tid] = incl - padded;
            qcur[tid] = incl - padded;
            qreal[tid] = c;
            qcnt[tid] = 0u;  // ready for the next tile (used again after two barriers)
            if (tid == 63u) seg[64] = incl;
        }
        if (MODE == AGG_CHUNK) retire_chunks();  // after the previous tile's last round
        lds_barrier();
        gather_a();
        bool late_gathers = true;  // B and C still to issue
        const uint32_t total = seg[64];
        // rounds over runs of whole query segments that fit the list (one
        // query's padded segment, at most R events, always fits). Lane q
        // holds the end of query q's segment; the ends do not decrease, so
        // the queries from qa on that end within ev_cap of the round's start
        // are a run, counted with one ballot
        const uint32_t seg_end = seg[lane + 1u];
        for (uint32_t qa = 0, lo = 0; qa < 64u && lo < total;) {
            const unsigned long long fit = __ballot(lane >= qa && seg_end - lo <= a.ev_cap);
            const uint32_t qb = max(qa + 1u, qa + (uint32_t)__popcll(fit));
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)seg_end, (int)(qb - 1u));
            const uint32_t n = hi - lo;  // a multiple of 64, at most ev_cap
            const unsigned long long qmask =
                (qb == 64u ? ~0ull : ((1ull << qb) - 1ull)) & ~((1ull << qa) - 1ull);
            // 4. list: a cursor atomic per match gives its slot in the
            // grouped order (row in bits 0..15, query in bits 16..21); the
            // padding slots of a segment hold row 0xffff
            for (int j = 0; j < 2; j++) {
                for (unsigned long long w = rotr64((j ? al1 : al0) & qmask, lane); w; w &= w - 1ull) {
                    const uint32_t q = ((uint32_t)__builtin_ctzll(w) + lane) & 63u;
                    const uint32_t pos = __hip_atomic_fetch_add(qcur + q, 1u, __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_WORKGROUP) - lo;
                    ev[pos] = (2u * tid + (uint32_t)j) | (q << 16);
                }
            }
            if (tid >= qa && tid < qb) {
                const uint32_t end = seg[tid] + ((qreal[tid] + 63u) & ~63u);
                for (uint32_t p = seg[tid] + qreal[tid]; p < end; p++) ev[p - lo] = 0xffffu | (tid << 16);
            }
            if (late_gathers) gather_b();
            lds_barrier();
            // 5. events: each wave takes whole 64-event chunks, all of one query
            const uint32_t chunks = n >> 6;
            // each chunk's event words are read one chunk ahead (the index
            // clamped to the last chunk: the read needs no branch)
            uint32_t e_next = ev[min(tid >> 6, chunks - 1u) * 64u + lane];
            for (uint32_t ch = tid >> 6; ch < chunks; ch += BS / 64u) {
                const uint32_t e = e_next;
                e_next = ev[min(ch + BS / 64u, chunks - 1u) * 64u + lane];
                const uint32_t q = __builtin_amdgcn_readfirstlane(e >> 16);
                const bool valid = (e & 0xffffu) != 0xffffu;
                const uint32_t row = valid ? (e & 0xffffu) : 0u;
                const c_desc &d = ((const c_desc *)a.desc)[q];
                batch_event<MODE>(a, d, valid, row, lds, rattr, rm, R, pc_cur, pc_spare, pc_fill, scur,
                                  sstart, CH, lane);
            }
            if (late_gathers) gather_c();
            late_gathers = false;
            lds_barrier();
            if (MODE == AGG_CHUNK && qb < 64u && hi < total) {  // another round follows
                retire_chunks();
                lds_barrier();
            }
            qa = qb;
            lo = hi;
        }
        if (late_gathers) {  // a tile without matches
            gather_b();
            gather_c();
        }
    }
    if (MODE == AGG_CHUNK) {
        retire_chunks();  // after the last tile's last round
        lds_barrier();
        // fills of the chunks still open (fill < CH)
        for (uint32_t p = tid; p < np; p += BS) {
            if (pc_cur[p] != kNoChunk) a.chunk.chunk_fill[pc_cur[p]] = pc_fill[p];
            if (pc_spare[p] != kNoChunk) a.chunk.chunk_fill[pc_spare[p]] = 0u;
        }
    }
    __syncthreads();
    if (MODE == AGG_COUNT) stage_end<MODE>(a.stage, lds);
    if (MODE == AGG_SCATTER) return;
    for (uint32_t q = 0; q < a.nq; q++) {
        const BQDesc &d = D[q];
        if (d.lds_tab == kNoLds) continue;
        const l_u64 *ls = (const l_u64 *)((l_u32 *)lds + d.lds_tab);
        const l_u32 *lc = (l_u32 *)lds + d.lds_tab + 2u * d.G;
        for (uint32_t g = tid; g < d.G; g += BS) {
            const uint32_t c = lc[g];
            if (c) {
                b_atomic_dev(&a.gcnt[d.out_off + g], (unsigned long long)c);
                b_atomic_dev(&a.gsum[d.out_off + g], ls[g]);
            }
        }
    }
}

// ------------------------------------------------------- segmented events
// Event-grouped scan without the count step. Every query owns a fixed
// segment of the tile's event list (16-bit row numbers), sized on the host
// from its expected matches per tile with room to spare, so a tile takes
// two barriers instead of four:
//  1. rows: as in k_batch_ev;
//  2. list: one LDS atomic per (row, query) match on the query's segment
//     word (its end in the high half, its cursor in the low half) places
//     the row in the segment; a match past the segment's end (more matches
//     in the tile than the segment holds) is aggregated at once by its lane
//     with global atomics (batch_event_direct);
//  3. events: every wave numbers the chunks of 64 placed events of all
//     queries itself (a prefix over the 64 segment fills) and takes every
//     NW-th chunk; the events of a chunk are one query's, as in k_batch_ev.
// The segment words alternate between two copies by tile: the copy of the
// next tile is reset during this tile's events step. One-pass staging only
// (AGG_CHUNK); the chunk retirement runs at the start of each list step,
// where no step touches the chunk state.

// a match of query q (per lane) aggregated directly into the outputs: the
// descriptor is read per lane, the row's attributes and measures from LDS
__device__ __forceinline__ void batch_event_direct(const BatchArgs &a, uint32_t q, uint32_t row, const l_u32 *rattr,
                                                const l_u32 *rm, uint32_t R)
{
    const BQDesc &d = a.desc[q];
    uint32_t gid = 0u;
#pragma unroll
    for (int t = 0; t < 4; t++)
        gid += (uint32_t)__umul64hi(d.M[t] * (uint64_t)rattr[4u * row + (uint32_t)t], (uint64_t)d.g[t]) * d.fac[t];
    if (gid >= d.G) {
        __hip_atomic_fetch_add(a.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const long long mv = (long long)(int32_t)rm[row] + (a.m2 ? (long long)(int32_t)rm[R + row] : 0ll);
    b_atomic_dev(&a.gsum[d.out_off + gid], (unsigned long long)mv);
    b_atomic_dev(&a.gcnt[d.out_off + gid], 1ull);
}

template <int BS, int FAST, bool CPT>
__global__ __launch_bounds__(BS) void k_batch_seg(BatchArgs a)
{
    constexpr int MODE = AGG_CHUNK;
    constexpr uint32_t R = 2u * BS;
    extern __shared__ __align__(16) uint32_t lds[];
    for (int t = 0; t < 4; t++) {
        if (a.mask_lds[t] != kNoLds) {
            const uint32_t *src = (const uint32_t *)a.gmask[t];
            for (uint32_t i = threadIdx.x; i < 2u * a.dim_rows[t]; i += BS) lds[a.mask_lds[t] + i] = src[i];
        }
        if (a.attr_lds[t] != kNoLds)
            for (uint32_t i = threadIdx.x; i < a.dim_rows[t]; i += BS) lds[a.attr_lds[t] + i] = a.grec[t][i].z;
    }
    for (uint32_t i = a.tab_begin + threadIdx.x; i < a.tab_end; i += BS) lds[i] = 0u;
    const uint32_t np = a.chunk.nparts;
    const uint32_t CH = 1u << a.chunk.ch_bits;
    stage_l_u32 *pc_cur = (stage_l_u32 *)lds + a.pc_lds;
    stage_l_u32 *pc_spare = pc_cur + np;
    stage_l_u32 *pc_fill = pc_spare + np;
    for (uint32_t p = threadIdx.x; p < np; p += BS) {
        pc_cur[p] = chunk_claim(a.chunk, p);
        pc_spare[p] = chunk_claim(a.chunk, p);
        pc_fill[p] = 0u;
    }
    l_u32 *L = (l_u32 *)lds;
    l_u32 *qseg = L + a.cnt_lds;       // [2][64] segment words: end << 16 | cursor
    l_u32 *sinit = qseg + 128u;        // [64] their value at the start of a tile
    l_u32 *rattr = L + a.row_lds;      // [R] the row's four attributes (16 bytes a row)
    l_u32 *rm = L + a.m_lds;           // [R] m1, then [R] m2
    typedef __attribute__((address_space(3))) uint16_t l_u16;
    l_u16 *ev = (l_u16 *)(L + a.ev_lds);
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    if (tid < 64u) {
        uint32_t w = 0u;
        if (tid < a.nq) {
            const uint32_t b = a.desc[tid].seg_base, c = a.desc[tid].seg_cap;
            w = ((b + c) << 16) | b;
        }
        sinit[tid] = w;
        qseg[tid] = w;
        qseg[64u + tid] = w;
    }
    __syncthreads();
    const uint64_t rows = a.rows;
    // the next tile's keys and measures, and its gathers of the records of
    // dimension rows outside LDS, are issued one tile ahead (as in
    // k_batch_ev): the gathers of each row's first member after the list
    // step, of its second after the events step
    const uint64_t rlast = rows - 2u;
    const int32_t *colp[6];
#pragma unroll
    for (int t = 0; t < 4; t++) colp[t] = (a.used >> t) & 1u ? a.fk[t] : a.m1;
    colp[4] = a.m1;
    colp[5] = a.m2 ? a.m2 : a.m1;
    int2 nk[4], nm1, nm2;
    uint4 ng[4][2];
    auto load_tile = [&](uint64_t tl) {
        const uint64_t r = tl * R + 2u * tid;
        const uint64_t rl = min(r, rlast);
        const bool shifted = r != rl;
        auto ld2 = [&](const int32_t *p) {
            const v2i_a4 v = __builtin_nontemporal_load((const v2i_a4 *)(p + rl));
            return make_int2(shifted ? v.y : v.x, v.y);
        };
#pragma unroll
        for (int t = 0; t < 4; t++) nk[t] = ld2(colp[t]);
        nm1 = ld2(colp[4]);
        nm2 = ld2(colp[5]);
    };
    auto gather = [&](int t, int j) {
        const bool hbm = ((a.used >> t) & 1u) && a.mask_lds[t] == kNoLds;
        const uint4 *g = hbm ? a.grec[t] : a.grec_any;
        const uint32_t kmax = hbm ? max(a.dim_rows[t], 1u) - 1u : 0u;
        ng[t][j] = g[min((uint32_t)(j ? nk[t].y : nk[t].x), kmax)];
    };
    auto gather_first = [&]() { gather(0, 0); gather(1, 0); gather(2, 0); gather(3, 0); };
    auto gather_second = [&]() { gather(0, 1); gather(1, 1); gather(2, 1); gather(3, 1); };
    if (blockIdx.x < a.ntiles) {
        load_tile(blockIdx.x);
        gather_first();
        gather_second();
    }
    auto retire_chunks = [&]() {
        for (uint32_t p = tid; p < np; p += BS) {
            const uint32_t f = pc_fill[p];
            if (f >= CH) {
                const uint32_t cur = pc_cur[p];
                if (cur != kNoChunk) a.chunk.chunk_fill[cur] = CH;
                pc_cur[p] = pc_spare[p];
                pc_fill[p] = f - CH;
                pc_spare[p] = chunk_claim(a.chunk, p);
            }
        }
    };
    uint32_t par = 0u;  // the copy of the segment words this tile uses
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t r0 = tile * R + 2u * tid;
        const uint64_t next = tile + gridDim.x < a.ntiles ? tile + gridDim.x : tile;
        int2 ck[4];
        uint4 cg[4][2];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            ck[t] = nk[t];
            cg[t][0] = ng[t][0];
            cg[t][1] = ng[t][1];
        }
        const int2 cm1 = nm1, cm2 = nm2;
        // 1. rows
        unsigned long long al0 = r0 < rows ? a.qvalid : 0ull;
        unsigned long long al1 = r0 + 1 < rows ? a.qvalid : 0ull;
        uint32_t atr[2][4] = {{0u, 0u, 0u, 0u}, {0u, 0u, 0u, 0u}};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            uint32_t at0 = 0u, at1 = 0u;
            if ((a.used >> t) & 1u) {
                const int2 k = ck[t];
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const uint32_t key = (uint32_t)(j ? k.y : k.x);
                    unsigned long long &al = j ? al1 : al0;
                    uint32_t &at = j ? at1 : at0;
                    if (al) {
                        if (key < a.dim_rows[t]) {
                            if (a.mask_lds[t] != kNoLds) {
                                const l_u64 *w = (const l_u64 *)(L + a.mask_lds[t]);
                                al &= w[key];
                                at = L[a.attr_lds[t] + key];
                            } else {
                                const uint4 rec = cg[t][j];
                                al &= (unsigned long long)rec.x | ((unsigned long long)rec.y << 32);
                                at = rec.z;
                            }
                        } else {
                            al &= a.nouse[t];
                        }
                    }
                }
                atr[0][t] = at0;
                atr[1][t] = at1;
            }
        }
        {
            typedef unsigned int v4u __attribute__((ext_vector_type(4)));
            typedef __attribute__((address_space(3))) v4u l_u4;
            l_u4 *r4 = (l_u4 *)rattr;
            r4[2u * tid] = v4u{atr[0][0], atr[0][1], atr[0][2], atr[0][3]};
            r4[2u * tid + 1u] = v4u{atr[1][0], atr[1][1], atr[1][2], atr[1][3]};
        }
        if (al0 | al1) {
            rm[2u * tid] = (uint32_t)cm1.x;
            rm[2u * tid + 1u] = (uint32_t)cm1.y;
            if (a.m2) {
                rm[R + 2u * tid] = (uint32_t)cm2.x;
                rm[R + 2u * tid + 1u] = (uint32_t)cm2.y;
            }
        }
        load_tile(next);
        retire_chunks();  // after the previous tile's events
        // 2. list: each lane walks its masks from its own lane number on (bit
        // b of the rotated mask is query (b + lane) mod 64), so the lanes of
        // one atomic instruction mostly hit different words; a match past
        // its segment's end writes the spare slot after the list and is
        // aggregated directly
        l_u32 *seg = qseg + 64u * par;
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const uint16_t row = (uint16_t)(2u * tid + (uint32_t)j);
            for (unsigned long long w = rotr64(j ? al1 : al0, lane); w; w &= w - 1ull) {
                const uint32_t q = ((uint32_t)__builtin_ctzll(w) + lane) & 63u;
                const uint32_t v = __hip_atomic_fetch_add(seg + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t pos = v & 0xffffu;
                const bool fits = pos < (v >> 16);
                ev[fits ? pos : a.ev_cap] = row;
                if (!fits) batch_event_direct(a, q, row, rattr, rm, R);
            }
        }
        gather_first();
        lds_barrier();
        // 3. events; wave 0 resets the other copy for the next tile. Every
        // wave numbers the chunks of 64 events of all queries itself (lane
        // q: query q's events placed in this tile, its chunks and the
        // exclusive sum of the chunks before it) and takes every NW-th
        // chunk, so the waves' shares differ by at most one chunk
        if (tid < 64u) qseg[64u * (par ^ 1u) + tid] = sinit[tid];
        const uint32_t sbase = sinit[lane] & 0xffffu;
        uint32_t nq_l, cstart, ctot;
        {
            const uint32_t v = seg[lane];
            nq_l = min(v & 0xffffu, v >> 16) - sbase;
            const uint32_t c = (nq_l + 63u) >> 6;
            uint32_t incl = c;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t u = (uint32_t)__shfl_up((int)incl, d);
                if (lane >= (uint32_t)d) incl += u;
            }
            cstart = incl - c;
            ctot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            nq_l = c ? nq_l : 0u;
        }
        // chunk i belongs to the last query whose chunks start at or before i
        // (queries without events take no chunk)
        auto owner = [&](uint32_t i) {
            const unsigned long long m = __ballot(cstart <= i && nq_l != 0u);
            return 63u - (uint32_t)__builtin_clzll(m);
        };
        constexpr uint32_t NW = BS / 64u;
        const uint32_t w0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));
        // each chunk's event words are read one chunk ahead (an index past
        // the last chunk reads the last one again)
        uint32_t i = w0;
        uint32_t qn = 0u, e_next = 0u;
        if (i < ctot) {
            qn = owner(i);
            const uint32_t k64 = (i - (uint32_t)__builtin_amdgcn_readlane((int)cstart, (int)qn)) << 6;
            e_next = ev[(uint32_t)__builtin_amdgcn_readlane((int)sbase, (int)qn) + k64 + lane];
        }
        for (; i < ctot; i += NW) {
            const uint32_t q = qn;
            const uint32_t k64 = (i - (uint32_t)__builtin_amdgcn_readlane((int)cstart, (int)q)) << 6;
            const uint32_t e = e_next;
            const uint32_t inext = min(i + NW, ctot - 1u);
            qn = owner(inext);
            e_next = ev[(uint32_t)__builtin_amdgcn_readlane((int)sbase, (int)qn) +
                        (((inext - (uint32_t)__builtin_amdgcn_readlane((int)cstart, (int)qn)) << 6)) + lane];
            const uint32_t n = (uint32_t)__builtin_amdgcn_readlane((int)nq_l, (int)q) - k64;
            const bool valid = lane < n;
            const uint32_t row = valid ? e : 0u;
            const c_desc &d = ((const c_desc *)a.desc)[q];
            batch_event<MODE, FAST, CPT>(a, d, valid, row, lds, rattr, rm, R, pc_cur, pc_spare, pc_fill, nullptr,
                                    nullptr, CH, lane);
        }
        gather_second();
        lds_barrier();
        par ^= 1u;
    }
    retire_chunks();  // after the last tile's events
    lds_barrier();
    for (uint32_t p = tid; p < np; p += BS) {
        if (pc_cur[p] != kNoChunk) a.chunk.chunk_fill[pc_cur[p]] = pc_fill[p];
        if (pc_spare[p] != kNoChunk) a.chunk.chunk_fill[pc_spare[p]] = 0u;
    }
    __syncthreads();
    for (uint32_t q = 0; q < a.nq; q++) {
        const BQDesc &d = a.desc[q];
        if (d.lds_tab == kNoLds) continue;
        const l_u64 *ls = (const l_u64 *)((l_u32 *)lds + d.lds_tab);
        const l_u32 *lc = (l_u32 *)lds + d.lds_tab + 2u * d.G;
        for (uint32_t g = tid; g < d.G; g += BS) {
            const uint32_t c = lc[g];
            if (c) {
                b_atomic_dev(&a.gcnt[d.out_off + g], (unsigned long long)c);
                b_atomic_dev(&a.gsum[d.out_off + g], ls[g]);
            }
        }
    }
}

// (threads, 128-bit chunks per thread per tile, LDS KiB for query tables,
// staged aggregation of the queries without an LDS table, event-grouped
// kernel with two rows per thread, copies of small tables in LDS)
struct BatchCfg {
    int bs;
    int unroll;
    int lds_kib;
    int staged;   // 0 atomics, 1 two-pass staging (count, scatter), 2 one-pass staging (chunks)
    int grouped;
    int small;
    int evk;   // grouped: event-list capacity in events per tile row (at most)
    int seg;   // grouped, one-pass: fixed per-query event segments (k_batch_seg)
};
const BatchCfg kBatchCfgs[] = {
    {512, 1, 48, 0, 0, 1, 0},   {512, 1, 96, 0, 0, 1, 0},   {256, 1, 48, 0, 0, 1, 0},
    {256, 1, 48, 1, 0, 1, 0},   {256, 1, 96, 1, 0, 1, 0},   {512, 1, 0, 1, 0, 1, 0},
    {512, 1, 48, 1, 0, 1, 0},   {512, 1, 96, 1, 0, 1, 0},   {512, 2, 48, 1, 0, 1, 0},
    {1024, 1, 48, 1, 0, 1, 0},  {1024, 1, 16, 1, 1, 1, 16}, {1024, 1, 16, 1, 1, 0, 16},
    {1024, 1, 0, 1, 1, 0, 4},   {1024, 1, 16, 0, 1, 1, 16}, {512, 1, 0, 1, 1, 0, 4},
    {512, 1, 0, 1, 1, 0, 2},    {512, 1, 16, 1, 1, 1, 16},  {512, 1, 16, 1, 1, 0, 4},
    {512, 1, 8, 1, 1, 0, 4},    {512, 1, 8, 2, 1, 0, 2},    {512, 1, 0, 2, 1, 0, 2},
    {1024, 1, 8, 2, 1, 0, 2},   {512, 1, 8, 2, 1, 0, 4},    {512, 1, 0, 2, 1, 0, 4},
    {512, 1, 0, 2, 1, 0, 3},    {512, 1, 8, 2, 1, 0, 6},    {512, 1, 8, 2, 1, 0, 8},
    {1024, 1, 8, 2, 1, 0, 4},   {512, 1, 8, 1, 1, 0, 8},    {512, 1, 16, 2, 1, 0, 4},
    {1024, 1, 0, 2, 1, 1, 6},   {1024, 1, 0, 2, 1, 1, 8},   {1024, 1, 4, 2, 1, 1, 6},
    {1024, 1, 0, 2, 1, 1, 0, 1}, {1024, 1, 4, 2, 1, 1, 0, 1}, {512, 1, 0, 2, 1, 0, 0, 1}};
const int kNumBatchCfgs = (int)(sizeof(kBatchCfgs) / sizeof(kBatchCfgs[0]));
// segmented kernel: at most kSegEvents events in a tile's list (16-bit
// positions that leave the segment word's cursor room to count past its end)
constexpr uint32_t kSegEvents = 32768;
// the mask kernels' pass counts (one per query and table) and a flag
constexpr uint32_t kPassWords = 4u * OLAP_BATCH_MAX + 1u;

// fast: the batch takes the 32-bit modulo; cpt: its pool holds 4-byte
// records (one measure), which some kernels are built for
template <int MODE>
static const void *batch_kernel(const BatchCfg &c, bool fast, bool cpt)
{
    if (c.seg) {
        if (MODE != AGG_CHUNK) return nullptr;
#define OLAP_SEG(BSX, F) \
    (cpt ? (const void *)k_batch_seg<BSX, F, true> : (const void *)k_batch_seg<BSX, F, false>)
        if (fast) return c.bs == 512 ? OLAP_SEG(512, 1) : OLAP_SEG(1024, 1);
        return c.bs == 512 ? OLAP_SEG(512, 0) : OLAP_SEG(1024, 0);
#undef OLAP_SEG
    }
    if (c.grouped) return c.bs == 512 ? (const void *)k_batch_ev<512, MODE> : (const void *)k_batch_ev<1024, MODE>;
    if (MODE == AGG_CHUNK) return nullptr;  // one-pass staging exists in the grouped kernel only
    if (c.bs == 256) return c.unroll == 1 ? (const void *)k_batch<256, 1, MODE> : (const void *)k_batch<256, 2, MODE>;
    if (c.bs == 512) return c.unroll == 1 ? (const void *)k_batch<512, 1, MODE> : (const void *)k_batch<512, 2, MODE>;
    return (const void *)k_batch<1024, 1, MODE>;
}

static const void *batch_kernel_mode(int mode, const BatchCfg &c, bool fast, bool cpt)
{
    if (mode == AGG_CHUNK) return batch_kernel<AGG_CHUNK>(c, fast, cpt);
    if (mode == AGG_COUNT) return batch_kernel<AGG_COUNT>(c, fast, cpt);
    if (mode == AGG_SCATTER) return batch_kernel<AGG_SCATTER>(c, fast, cpt);
    return batch_kernel<AGG_DIRECT>(c, fast, cpt);
}

// fact rows per tile of a configuration
static uint64_t batch_tile_rows(const BatchCfg &c)
{
    return c.grouped ? 2ull * (uint64_t)c.bs : (uint64_t)c.bs * 4u * (uint64_t)c.unroll;
}

// grid of a configuration: persistent, limited by the least occupant of the
// kernels it runs (COUNT and SCATTER of a staged run share one grid)
static hipError_t batch_grid(int cfg, uint32_t dyn, int num_cus, uint64_t rows, bool fast, bool cpt, int *grid)
{
    const BatchCfg c = kBatchCfgs[cfg];
    const int modes[2] = {c.staged == 2 ? AGG_CHUNK : c.staged ? AGG_COUNT : AGG_DIRECT,
                          c.staged == 2 ? AGG_CHUNK : c.staged ? AGG_SCATTER : AGG_DIRECT};
    int best = 1 << 30;
    for (int m : modes) {
        const void *k = batch_kernel_mode(m, c, fast, cpt);
        if (dyn > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
            if (e != hipSuccess) return e;
        }
        int per_cu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, c.bs, dyn);
        if (e != hipSuccess) return e;
        best = std::min(best, per_cu);
    }
    if (best < 1) return hipErrorInvalidConfiguration;
    const uint64_t tile = batch_tile_rows(c);
    const uint64_t tiles = (rows + tile - 1) / tile;
    *grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)best * num_cus, tiles));
    return hipSuccess;
}

static hipError_t batch_launch(int cfg, int mode, BatchArgs &a, uint32_t dyn, int grid,
                               hipStream_t s)
{
    const BatchCfg c = kBatchCfgs[cfg];
    const uint64_t tile = batch_tile_rows(c);
    a.ntiles = (a.rows + tile - 1) / tile;
    if (a.ntiles == 0) return hipSuccess;
    void *args[] = {&a};
    return hipLaunchKernel(batch_kernel_mode(mode, c, a.fastmod != 0u, a.chunk.fmt == REC_COMPACT), dim3(grid),
                           dim3(c.bs), args, dyn, s);
}

}  // namespace olap

namespace olap {

struct BatchWorkspace {
    Stager stager;
    unsigned long long *mask[4] = {nullptr, nullptr, nullptr, nullptr};  // [dim rows] per table
    uint4 *rec[4] = {nullptr, nullptr, nullptr, nullptr};                 // [dim rows] per table
    size_t dim_cap[4] = {0, 0, 0, 0};
    BQDesc *desc[kNumBatchCfgs] = {};   // OLAP_BATCH_MAX descriptors per configuration
    MaskPred *preds = nullptr;          // up to 4 x OLAP_BATCH_MAX predicates (one per query and table)
    uint32_t *pass = nullptr;           // their pass counts, then the negative-attribute flag
    int32_t *minmax = nullptr;          // [2] device scratch of measure_range
    int8_t *progvec = nullptr;          // pass vectors of the steps' general predicates
    size_t progvec_cap = 0;
    // the ranges of the last kRanges measure columns seen (a server worker
    // alternates between datasets): column, rows, min and max
    static constexpr int kRanges = 8;
    struct Range {
        const int32_t *col = nullptr;
        int64_t rows = -1;
        int32_t lo = 0, hi = 0;
    } ranges[kRanges];
    int next_range = 0;

    hipError_t fit_table(int t, size_t rows)
    {
        rows = std::max<size_t>(rows, 1);
        if (mask[t] && rows <= dim_cap[t]) return hipSuccess;
        if (mask[t]) (void)hipFree(mask[t]);
        if (rec[t]) (void)hipFree(rec[t]);
        mask[t] = nullptr;
        rec[t] = nullptr;
        dim_cap[t] = 0;
        hipError_t e = hipMalloc(&mask[t], 8 * rows);
        if (e == hipSuccess) e = hipMalloc(&rec[t], 16 * rows);
        if (e == hipSuccess) dim_cap[t] = rows;
        return e;
    }
    hipError_t fit_progvec(size_t bytes)
    {
        if (bytes <= progvec_cap) return hipSuccess;
        if (progvec) (void)hipFree(progvec);
        progvec = nullptr;
        progvec_cap = 0;
        hipError_t e = hipMalloc(&progvec, bytes);
        if (e == hipSuccess) progvec_cap = bytes;
        return e;
    }
    hipError_t fit_fixed()
    {
        hipError_t e = hipSuccess;
        for (int c = 0; c < kNumBatchCfgs && e == hipSuccess; c++)
            if (!desc[c]) e = hipMalloc(&desc[c], sizeof(BQDesc) * OLAP_BATCH_MAX);
        if (e == hipSuccess && !preds) e = hipMalloc(&preds, sizeof(MaskPred) * 4 * OLAP_BATCH_MAX);
        if (e == hipSuccess && !pass) e = hipMalloc(&pass, 4 * kPassWords);
        if (e == hipSuccess && !minmax) e = hipMalloc(&minmax, 8);
        return e;
    }
    ~BatchWorkspace()
    {
        for (int t = 0; t < 4; t++) {
            if (mask[t]) (void)hipFree(mask[t]);
            if (rec[t]) (void)hipFree(rec[t]);
        }
        for (int c = 0; c < kNumBatchCfgs; c++)
            if (desc[c]) (void)hipFree(desc[c]);
        if (preds) (void)hipFree(preds);
        if (pass) (void)hipFree(pass);
        if (minmax) (void)hipFree(minmax);
        if (progvec) (void)hipFree(progvec);
    }
};

BatchWorkspace *batch_workspace_create() { return new BatchWorkspace(); }
void batch_workspace_destroy(BatchWorkspace *ws) { delete ws; }

}  // namespace olap

using namespace olap;

struct olap_batch_s {
    olap_ctx ctx = nullptr;
    olap_star star{};
    int nq = 0;
    std::vector<olap_batch_query> queries;
    std::vector<int64_t> offsets;
    BatchWorkspace *ws = nullptr;       // masks, descriptors, staging scratch
    bool ws_owned = false;
    hipStream_t setup_s = nullptr;      // stream of mask builds and descriptor uploads
    std::vector<double> sel;
    double events = 0.0;
    uint32_t used = 0;
    unsigned long long nouse[4] = {0, 0, 0, 0};
    int cfg = 6;
    // per-config launch plan
    BatchArgs args[kNumBatchCfgs];
    uint32_t dyn[kNumBatchCfgs] = {};
    int grid[kNumBatchCfgs] = {};
    uint32_t stage_lds[kNumBatchCfgs] = {};  // word offset of the stage region
    bool planned[kNumBatchCfgs] = {};
    // the workspace's staging scratch is set up for one configuration's grid
    int stager_cfg = -1;
    std::vector<unsigned long long> out_off;
    std::vector<uint32_t> gsz;
    uint32_t nparts = 0;
    uint32_t nparts_c = 0;   // partitions of 2^12 slots (compact one-pass records)
    bool compact = false;    // one measure whose range spans fewer than 2^20 values
    bool fastmod = false;    // every probed table's attributes in [0, 2^31), every G below 2^24
    int32_t mbase = 0;       // its minimum
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

olap_status bfail(olap_ctx ctx, olap_status st, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return st;
}

olap_status bhip(olap_ctx ctx, hipError_t e, const char *what)
{
    return bfail(ctx, e == hipErrorOutOfMemory ? OLAP_E_NOMEM : OLAP_E_HIP,
                 std::string(what) + ": " + hipGetErrorString(e));
}

uint32_t up16(uint32_t bytes) { return (bytes + 15u) & ~15u; }

// the multiplier and shift of fast_term for modulo g (0 and 0 for g 0 or 1)
void fastmod_magic(uint32_t g, uint32_t *m, uint32_t *sh)
{
    *m = 0u;
    *sh = 0u;
    if (g < 2u) return;
    const uint32_t s = 31u - (uint32_t)__builtin_clz(g);
    if ((g & (g - 1u)) == 0u) {
        *m = 1u << 31;
        *sh = s - 1u;
    } else {
        *m = (uint32_t)(((1ull << (32u + s)) + g - 1u) / g);
        *sh = s;
    }
}

// Segments of the segmented kernel's event list: query q expects
// E = R * sel[q] matches a tile; its segment holds E + 4 sqrt(E) rounded up
// to whole chunks of 64 (at least one), scaled down to the list when the sum
// exceeds it; the chunks left over go to the queries in turn, most matches
// first, so skewed data leaves fewer residual passes.
void plan_segments(olap_batch b, const BatchCfg &cfg, uint32_t cap_events, std::vector<BQDesc> &desc)
{
    const int nq = b->nq;
    const double R = (double)batch_tile_rows(cfg);
    const uint32_t chunks = cap_events / 64u;   // >= nq (ev_min)
    std::vector<uint32_t> k((size_t)nq);
    uint64_t total = 0;
    for (int q = 0; q < nq; q++) {
        const double E = std::min(R, std::max(0.0, b->sel[(size_t)q]) * R);
        k[(size_t)q] = std::max<uint32_t>(1u, (uint32_t)std::ceil((E + 4.0 * std::sqrt(E)) / 64.0));
        total += k[(size_t)q];
    }
    if (total > chunks) {
        uint64_t t2 = 0;
        for (int q = 0; q < nq; q++) {
            k[(size_t)q] = std::max<uint32_t>(1u, (uint32_t)((uint64_t)k[(size_t)q] * (chunks - (uint32_t)nq) / total));
            t2 += k[(size_t)q];
        }
        total = t2;  // sum of max(1, floor(k (chunks - nq) / total)) <= chunks
    }
    std::vector<int> order((size_t)nq);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return b->sel[(size_t)x] > b->sel[(size_t)y]; });
    // a segment never needs more than a tile's rows
    const uint32_t kmax = (uint32_t)(R / 64.0);
    for (bool grew = true; total < chunks && grew;) {
        grew = false;
        for (int q : order) {
            if (total >= chunks) break;
            if (k[(size_t)q] < kmax) {
                k[(size_t)q]++;
                total++;
                grew = true;
            }
        }
    }
    uint32_t base = 0;
    for (int q = 0; q < nq; q++) {
        desc[(size_t)q].seg_base = base;
        desc[(size_t)q].seg_cap = 64u * k[(size_t)q];
        base += 64u * k[(size_t)q];
    }
}

// LDS layout of one configuration: descriptors, then masks/attributes of the
// smallest tables, then per-query group tables by events per byte.
olap_status plan_config(olap_batch b, int c)
{
    olap_ctx ctx = b->ctx;
    const BatchCfg cfg = kBatchCfgs[c];
    BatchArgs &a = b->args[c];
    memset(&a, 0, sizeof(a));
    std::vector<BQDesc> desc((size_t)b->nq);
    uint32_t cur = 0;
    a.desc_lds = 0;
    // the row-wise batch kernel reads the descriptors from LDS, the
    // event-grouped one from global memory (wave-uniform scalar loads)
    if (!cfg.grouped) cur += up16((uint32_t)(sizeof(BQDesc) * b->nq));
    const uint32_t table_budget = (uint32_t)cfg.lds_kib * 1024u;
    // staged configurations keep a u32 counter and a u64 range start per partition
    const uint32_t stage_bytes = cfg.staged == 1 ? up16(b->nparts * 4u) + b->nparts * 8u
                                 : cfg.staged == 2 ? up16((b->compact ? b->nparts_c : b->nparts) * 12u) : 0u;
    // the event-grouped kernel stages per-tile rows and at least one event
    // per row of the tile
    const uint32_t R = (uint32_t)batch_tile_rows(cfg);
    // (the segmented kernel: 16-bit events, at least 64 per query)
    const uint32_t ev_bytes = cfg.seg ? 2u : 4u;
    const uint32_t ev_min = cfg.seg ? ev_bytes * 64u * (uint32_t)b->nq : cfg.grouped ? ev_bytes * (R + 64u) : 0u;
    if (cfg.grouped) {
        a.cnt_lds = cur / 4u;
        // counters, cursors, starts and the total; or the two copies of the
        // segment words and their initial values
        cur += cfg.seg ? up16(4u * (128u + 64u)) : up16(4u * (64u + 64u + 64u + 65u));
        a.row_lds = cur / 4u;
        cur += 16u * R;
        a.m_lds = cur / 4u;
        cur += (b->star.m2 ? 8u : 4u) * R;
    }
    // the event-grouped kernel reads the fact columns in row pairs with no
    // tail branch (the last pair in bounds stands in for pairs past the end)
    if (cfg.grouped && b->star.rows < 2)
        return bfail(ctx, OLAP_E_CAPACITY, "batch config " + std::to_string(c) + ": needs at least 2 fact rows");
    if ((uint64_t)cur + stage_bytes + ev_min + 1024u > kLdsBytes)
        return bfail(ctx, OLAP_E_CAPACITY, "batch config " + std::to_string(c) + ": LDS too small for " +
                                               std::to_string(b->nparts) + " partitions");
    const uint32_t cap = kLdsBytes - 1024u - stage_bytes - ev_min;
    // masks + attributes of probed tables, smallest first, within cap - table budget
    std::vector<int> order;
    for (int t = 0; t < 4; t++)
        if ((b->used >> t) & 1u) order.push_back(t);
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return b->star.dim_rows[x] < b->star.dim_rows[y]; });
    for (int t = 0; t < 4; t++) {
        a.mask_lds[t] = kNoLds;
        a.attr_lds[t] = kNoLds;
    }
    for (int t : order) {
        const uint32_t need = up16(8u * (uint32_t)b->star.dim_rows[t]) + up16(4u * (uint32_t)b->star.dim_rows[t]);
        if (cfg.small && cur + need + table_budget <= cap) {
            a.mask_lds[t] = cur / 4u;
            cur += up16(8u * (uint32_t)b->star.dim_rows[t]);
            a.attr_lds[t] = cur / 4u;
            cur += up16(4u * (uint32_t)b->star.dim_rows[t]);
        }
    }
    // the kernels' numbering of the tables, pm[t]: the probed tables outside
    // LDS take positions 0, 2, 1, 3 in turn (the one-pass form spreads its
    // gathers by position), the others the positions left
    int pm[4];
    {
        static const int kHbmPos[4] = {0, 2, 1, 3};
        bool taken[4] = {false, false, false, false};
        auto hbm = [&](int t) { return ((b->used >> t) & 1u) && a.mask_lds[t] == kNoLds; };
        int n = 0;
        for (int t = 0; t < 4; t++)
            if (hbm(t)) {
                pm[t] = kHbmPos[n++];
                taken[pm[t]] = true;
            }
        int next = 0;
        for (int t = 0; t < 4; t++) {
            if (hbm(t)) continue;
            while (taken[next]) next++;
            pm[t] = next;
            taken[next] = true;
        }
        uint32_t ml[4], al[4];
        for (int t = 0; t < 4; t++) {
            ml[pm[t]] = a.mask_lds[t];
            al[pm[t]] = a.attr_lds[t];
        }
        for (int t = 0; t < 4; t++) {
            a.mask_lds[t] = ml[t];
            a.attr_lds[t] = al[t];
        }
    }
    a.tab_begin = cur / 4u;
    // query tables: highest expected events per byte first
    std::vector<int> qs((size_t)b->nq);
    std::iota(qs.begin(), qs.end(), 0);
    std::sort(qs.begin(), qs.end(), [&](int x, int y) {
        return b->sel[x] / (double)b->queries[x].group_nums > b->sel[y] / (double)b->queries[y].group_nums;
    });
    uint32_t tab_used = 0;
    for (int q : qs) desc[q].lds_tab = kNoLds;
    for (int q : qs) {
        const uint64_t need = up16((uint32_t)std::min<int64_t>(b->queries[q].group_nums * 12, 1u << 30));
        if (b->queries[q].group_nums * 12 <= (int64_t)table_budget && tab_used + need <= table_budget &&
            cur + tab_used + need <= cap) {
            desc[q].lds_tab = (cur + tab_used) / 4u;
            tab_used += (uint32_t)need;
        }
    }
    cur += tab_used;
    a.tab_end = cur / 4u;
    b->stage_lds[c] = cur / 4u;
    cur += stage_bytes;
    if (cfg.staged == 2) a.pc_lds = b->stage_lds[c];
    if (cfg.seg) {
        // the rest of LDS holds the event segments
        a.ev_cap = std::min<uint32_t>(kSegEvents, (kLdsBytes - 1024u - 16u - cur) / ev_bytes) & ~63u;
        a.ev_lds = cur / 4u;
        cur += up16(ev_bytes * (a.ev_cap + 1u));  // (+ the spare slot of matches past a segment)
    } else if (cfg.grouped) {
        // the rest of LDS holds the event list (at most evk events per row)
        a.ev_cap = std::min<uint32_t>(std::max<uint32_t>((uint32_t)cfg.evk * R, R + 64u),
                                      (kLdsBytes - 1024u - cur) / ev_bytes) & ~63u;
        a.ev_lds = cur / 4u;
        cur += 4u * a.ev_cap;
    }
    if (cfg.seg) plan_segments(b, cfg, a.ev_cap, desc);
    uint32_t pbase = 0, pbase_c = 0;
    for (int q = 0; q < b->nq; q++) {
        const olap_batch_query &Q = b->queries[q];
        BQDesc &d = desc[q];
        d.nsteps = (uint32_t)Q.ndims;
        d.G = (uint32_t)Q.group_nums;
        d.out_off = (uint32_t)b->offsets[q];
        d.pbase = pbase;
        d.pbase1 = b->compact ? pbase_c : pbase;
        pbase += (d.G + kPartSlots - 1) / kPartSlots;
        pbase_c += (d.G + (1u << kCompactPartBits) - 1) >> kCompactPartBits;
        for (int s = 0; s < Q.ndims; s++) {
            const int t = pm[Q.table[s]];
            d.g[t] = (uint32_t)Q.group_mod[s];
            d.fac[t] = d.g[t] > 1u ? (uint32_t)Q.factor[s] : 0u;  // modulo 0 or 1: code 0
            d.M[t] = d.g[t] ? (UINT64_MAX / d.g[t] + 1ull) : 0ull;
            uint32_t m = 0u, sh = 0u;
            fastmod_magic(d.g[t], &m, &sh);
            d.m32[t] = m;
            d.sh4 |= sh << (8 * t);
        }
    }
    hipError_t e = hipMemcpyAsync(b->ws->desc[c], desc.data(), sizeof(BQDesc) * b->nq, hipMemcpyHostToDevice,
                                  b->setup_s);
    if (e == hipSuccess) e = hipStreamSynchronize(b->setup_s);  // desc is read by the copy
    if (e != hipSuccess) return bhip(ctx, e, "batch descriptors");
    a.used = 0u;
    for (int t = 0; t < 4; t++) {
        const int k = pm[t];
        a.fk[k] = b->star.fk[t];
        a.gmask[k] = b->ws->mask[t];
        a.grec[k] = b->ws->rec[t];
        a.dim_rows[k] = (uint32_t)std::max<int64_t>(b->star.dim_rows[t], 0);
        a.nouse[k] = b->nouse[t];
        if ((b->used >> t) & 1u) {
            a.used |= 1u << k;
            if (!a.grec_any) a.grec_any = b->ws->rec[t];
        }
    }
    a.m1 = b->star.m1;
    a.m2 = b->star.m2;
    a.rows = (uint64_t)b->star.rows;
    a.nq = (uint32_t)b->nq;
    a.qvalid = b->nq >= 64 ? ~0ull : (1ull << b->nq) - 1ull;
    a.fastmod = b->fastmod ? 1u : 0u;
    a.desc = b->ws->desc[c];
    a.err = ctx->d_err;
    b->dyn[c] = std::max<uint32_t>(cur, 16u);
    e = batch_grid(c, b->dyn[c], ctx->num_cus, a.rows, a.fastmod != 0u, !b->star.m2 && b->compact, &b->grid[c]);
    if (e != hipSuccess) return bhip(ctx, e, "batch launch plan");
    b->planned[c] = true;
    return OLAP_OK;
}

olap_status stager_for(olap_batch b, int c, hipStream_t s)
{
    if (b->stager_cfg == c) return OLAP_OK;
    const uint32_t pbits = kBatchCfgs[c].staged == 2 && b->compact ? kCompactPartBits : kPartBits;
    hipError_t e = b->ws->stager.setup(b->out_off.data(), b->gsz.data(), b->nq, (uint32_t)b->grid[c],
                                       b->star.m2 != nullptr, s, pbits);
    if (e != hipSuccess) return bhip(b->ctx, e, "stager setup");
    b->stager_cfg = c;
    return OLAP_OK;
}

// zero + scan(s) + reduce of the current configuration; `sized` skips the
// host read of the record total (graph capture: the buffers were sized by an
// earlier run of the same batch)
olap_status batch_run(olap_batch b, int64_t *sum, int64_t *count, hipStream_t s, bool sized)
{
    olap_ctx ctx = b->ctx;
    const int c = b->cfg;
    const size_t slots = (size_t)b->offsets[b->nq];
    hipError_t e = launch_zero_u64((unsigned long long *)sum, slots, s);
    if (e == hipSuccess) e = launch_zero_u64((unsigned long long *)count, slots, s);
    if (e != hipSuccess) return bhip(ctx, e, "batch zero");
    BatchArgs a = b->args[c];
    a.gsum = (unsigned long long *)sum;
    a.gcnt = (unsigned long long *)count;
    if (!kBatchCfgs[c].staged) {
        e = batch_launch(c, AGG_DIRECT, a, b->dyn[c], b->grid[c], s);
        return e == hipSuccess ? OLAP_OK : bhip(ctx, e, "batch scan");
    }
    olap_status st = stager_for(b, c, s);
    if (st != OLAP_OK) return st;
    Stager &S = b->ws->stager;
    if (kBatchCfgs[c].staged == 2) {
        // one pass: the pool is sized from the expected matches (it only
        // grows, so a captured replay never reallocates); matches beyond it
        // go to global atomics inside the scan
        const uint32_t ch_bits = (uint32_t)__builtin_ctz((uint32_t)batch_tile_rows(kBatchCfgs[c]));
        if (!sized) {
            const double expect = std::max(0.0, b->events) * (double)b->star.rows;
            const uint32_t fmt = b->star.m2 ? REC_WIDE : b->compact ? REC_COMPACT : REC_NARROW;
            e = S.chunk_reserve((unsigned long long)expect, (uint32_t)b->grid[c], ch_bits, fmt, b->mbase);
            if (e != hipSuccess) return bhip(ctx, e, "batch chunk pool");
        }
        e = S.chunk_clear(s);
        a.chunk = S.chunk_args();
        if (e == hipSuccess) e = batch_launch(c, AGG_CHUNK, a, b->dyn[c], b->grid[c], s);
        if (e == hipSuccess) e = S.chunk_reduce((unsigned long long *)sum, (unsigned long long *)count,
                                                ctx->num_cus, s);
        return e == hipSuccess ? OLAP_OK : bhip(ctx, e, "batch one-pass scan/reduce");
    }
    const uint32_t cur = b->stage_lds[c];
    const uint32_t start = cur + up16(b->nparts * 4u) / 4u;
    a.stage = S.args(cur, start, a.gsum, a.gcnt);
    e = batch_launch(c, AGG_COUNT, a, b->dyn[c], b->grid[c], s);
    if (e == hipSuccess) e = S.prefix(s);
    if (e != hipSuccess) return bhip(ctx, e, "batch count pass");
    if (!sized || !S.rec) {
        unsigned long long total = 0;
        e = S.total(&total, s);
        if (e == hipSuccess) e = S.reserve(total);
        if (e != hipSuccess) return bhip(ctx, e, "batch staging buffers");
        a.stage = S.args(cur, start, a.gsum, a.gcnt);
    }
    e = batch_launch(c, AGG_SCATTER, a, b->dyn[c], b->grid[c], s);
    if (e == hipSuccess) e = S.reduce((unsigned long long *)sum, (unsigned long long *)count,
                                      ctx->num_cus, s);
    return e == hipSuccess ? OLAP_OK : bhip(ctx, e, "batch scatter/reduce");
}

}  // namespace

olap_status olap::batch_create_in(olap_ctx ctx, const olap_star *star, const olap_batch_query *queries,
                                  int32_t nq, BatchWorkspace *ws, hipStream_t strm, olap_batch *out)
{
    if (!ctx) return OLAP_E_INVALID;
    if (!star || !queries || !out || !ws || nq < 1 || nq > OLAP_BATCH_MAX)
        return bfail(ctx, OLAP_E_INVALID, "batch_create: 1..64 queries and a star are required");
    *out = nullptr;
    if (star->rows < 0 || star->rows > (int64_t)UINT32_MAX - 65536)
        return bfail(ctx, OLAP_E_CAPACITY, "batch_create: rows out of range");
    if (!star->m1 && star->rows) return bfail(ctx, OLAP_E_INVALID, "batch_create: m1 is required");
    auto aligned = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    if (!aligned(star->m1) || !aligned(star->m2))
        return bfail(ctx, OLAP_E_INVALID, "batch_create: fact columns must be 16-byte aligned");
    olap_batch b = new olap_batch_s();
    b->ctx = ctx;
    b->ws = ws;
    b->setup_s = strm;
    b->star = *star;
    b->nq = nq;
    b->queries.assign(queries, queries + nq);
    b->offsets.assign((size_t)nq + 1, 0);
    for (int q = 0; q < nq; q++) {
        const olap_batch_query &Q = queries[q];
        std::string where = "query " + std::to_string(q) + ": ";
        if (Q.ndims < 1 || Q.ndims > 4 || Q.group_nums < 1 || Q.group_nums > INT32_MAX) {
            delete b;
            return bfail(ctx, OLAP_E_INVALID, where + "ndims 1..4 and group_nums in [1, 2^31) required");
        }
        int64_t max_gid = 0;
        uint32_t seen = 0;
        for (int s = 0; s < Q.ndims; s++) {
            const int t = Q.table[s];
            if (t < 0 || t > 3 || ((seen >> t) & 1u) || Q.group_mod[s] < 0 || Q.factor[s] < 0) {
                delete b;
                return bfail(ctx, OLAP_E_INVALID, where + "distinct tables 0..3, group_mod >= 0, factor >= 0");
            }
            if (!star->filter[t] || !star->fk[t] || (Q.group_mod[s] > 0 && !star->attr[t]) ||
                star->dim_rows[t] < 0 || star->dim_rows[t] > INT32_MAX || !aligned(star->fk[t])) {
                delete b;
                return bfail(ctx, OLAP_E_INVALID, where + "table " + std::to_string(t) +
                                                      " lacks filter/attr/fk columns (fk 16-byte aligned)");
            }
            seen |= 1u << t;
            if (Q.group_mod[s] > 1) max_gid += (int64_t)(Q.group_mod[s] - 1) * Q.factor[s];
            b->used |= 1u << t;
        }
        if (max_gid >= Q.group_nums) {
            delete b;
            return bfail(ctx, OLAP_E_INVALID, where + "largest gid reaches group_nums");
        }
        b->offsets[q + 1] = b->offsets[q] + Q.group_nums;
    }
    if (b->offsets[nq] > INT32_MAX) {
        delete b;
        return bfail(ctx, OLAP_E_CAPACITY, "batch_create: more than 2^31 output slots in total");
    }
    for (int t = 0; t < 4; t++) {
        b->nouse[t] = 0;
        for (int q = 0; q < nq; q++) {
            bool uses = false;
            for (int s = 0; s < queries[q].ndims; s++) uses |= queries[q].table[s] == t;
            if (!uses) b->nouse[t] |= 1ull << q;
        }
    }
    // masks and their pass statistics, on the setup stream: the predicates
    // of every probed table in one upload, one mask kernel per table, one
    // read-back of the pass counts and one synchronisation
    b->sel.assign((size_t)nq, 1.0);
    hipError_t e = ws->fit_fixed();
    // general predicates: each step's program is evaluated once into an
    // int8 pass vector of its table, which the mask kernel ANDs in
    size_t prog_bytes = 0;
    for (int q = 0; q < nq; q++)
        for (int s = 0; s < queries[q].ndims; s++)
            if (queries[q].program[s]) prog_bytes += ((size_t)star->dim_rows[queries[q].table[s]] + 15) & ~(size_t)15;
    if (e == hipSuccess && prog_bytes) e = ws->fit_progvec(prog_bytes);
    if (e != hipSuccess) {
        delete b;
        return bhip(ctx, e, "batch predicates");
    }
    std::vector<MaskPred> preds;
    std::vector<int> pq;
    uint32_t pbeg[5] = {0, 0, 0, 0, 0};
    size_t poff = 0;
    for (int t = 0; t < 4; t++) {
        pbeg[t] = (uint32_t)preds.size();
        if ((b->used >> t) & 1u)
            for (int q = 0; q < nq; q++)
                for (int s = 0; s < queries[q].ndims; s++)
                    if (queries[q].table[s] == t) {
                        const int8_t *pv = nullptr;
                        if (const olap_dim_program *prog = queries[q].program[s]) {
                            DevProgram dp;
                            olap_status pst = prepare_program(ctx, prog, star->dim_rows[t], &dp, strm);
                            if (pst != OLAP_OK) {
                                delete b;
                                return pst;
                            }
                            dp.group = OLAP_GROUP_NONE;  // only the pass matters here
                            e = launch_dim_program(dp, star->dim_rows[t], 1, ws->progvec + poff,
                                                   (uint32_t *)ctx->d_scratch, strm);
                            pv = ws->progvec + poff;
                            poff += ((size_t)star->dim_rows[t] + 15) & ~(size_t)15;
                        }
                        preds.push_back({q, queries[q].lo[s], queries[q].hi[s], 0, pv});
                        pq.push_back(q);
                    }
        if (e == hipSuccess && ((b->used >> t) & 1u)) e = ws->fit_table(t, (size_t)star->dim_rows[t]);
    }
    pbeg[4] = (uint32_t)preds.size();
    if (e == hipSuccess)
        e = hipMemcpyAsync(ws->preds, preds.data(), sizeof(MaskPred) * preds.size(), hipMemcpyHostToDevice, strm);
    if (e == hipSuccess) e = launch_zero_u32(ws->pass, kPassWords, strm);
    for (int t = 0; t < 4 && e == hipSuccess; t++) {
        const uint32_t D = (uint32_t)star->dim_rows[t];
        if (!((b->used >> t) & 1u) || !D) continue;
        const int grid = (int)std::min<uint32_t>((D + kMaskThreads - 1) / kMaskThreads, 2u * (uint32_t)ctx->num_cus);
        k_batch_masks<<<grid, kMaskThreads, 0, strm>>>(star->filter[t], star->attr[t], D, ws->preds + pbeg[t],
                                              (int)(pbeg[t + 1] - pbeg[t]), b->nouse[t], ws->mask[t], ws->rec[t],
                                              ws->pass + pbeg[t], ws->pass + kPassWords - 1);
        e = hipGetLastError();
    }
    std::vector<uint32_t> pass(kPassWords, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(pass.data(), ws->pass, 4 * kPassWords, hipMemcpyDeviceToHost, strm);
    if (e == hipSuccess) e = hipStreamSynchronize(strm);
    for (int t = 0; t < 4 && e == hipSuccess; t++) {
        const uint32_t D = (uint32_t)star->dim_rows[t];
        for (uint32_t i = pbeg[t]; i < pbeg[t + 1]; i++)
            b->sel[(size_t)pq[i]] *= D ? (double)pass[i] / (double)D : 0.0;
    }
    if (e != hipSuccess) {
        delete b;
        return bhip(ctx, e, "batch masks");
    }
    // one measure whose values span fewer than 2^20: the one-pass records
    // are 4 bytes (the range is kept per workspace and column; a measure
    // outside it later, the columns having changed, is added directly)
    if (!star->m2 && star->rows > 0) {
        BatchWorkspace::Range *hit = nullptr;
        for (BatchWorkspace::Range &r : ws->ranges)
            if (r.col == star->m1 && r.rows == star->rows) hit = &r;
        if (!hit) {
            int32_t mm[2] = {0, 0};
            e = measure_range(star->m1, (uint64_t)star->rows, ws->minmax, mm, strm);
            if (e != hipSuccess) {
                delete b;
                return bhip(ctx, e, "batch measure range");
            }
            hit = &ws->ranges[ws->next_range];
            ws->next_range = (ws->next_range + 1) % BatchWorkspace::kRanges;
            *hit = {star->m1, star->rows, mm[0], mm[1]};
        }
        b->compact = (int64_t)hit->hi - (int64_t)hit->lo < (1ll << kCompactMeasureBits);
        b->mbase = hit->lo;
    }
    // the 32-bit modulo of the group-by terms: every attribute of the probed
    // tables non-negative and every G below 2^24
    b->fastmod = e == hipSuccess && pass[kPassWords - 1] == 0u;
    for (int q = 0; q < nq; q++) b->fastmod = b->fastmod && queries[q].group_nums < (1 << 24);
    b->events = 0;
    for (double v : b->sel) b->events += v;
    for (int q = 0; q < nq; q++) {
        b->out_off.push_back((unsigned long long)b->offsets[q]);
        b->gsz.push_back((uint32_t)queries[q].group_nums);
        b->nparts += (uint32_t)((queries[q].group_nums + kPartSlots - 1) / kPartSlots);
        b->nparts_c += (uint32_t)((queries[q].group_nums + (1 << kCompactPartBits) - 1) >> kCompactPartBits);
    }
    if (b->nparts_c > kMaxParts) b->compact = false;
    if (b->nparts > kMaxParts) {
        delete b;
        return bfail(ctx, OLAP_E_CAPACITY, "batch_create: too many aggregation partitions");
    }
    olap_status st = plan_config(b, b->cfg);
    if (st != OLAP_OK) {
        delete b;
        return st;
    }
    *out = b;
    return OLAP_OK;
}

extern "C" {

olap_status olap_batch_create(olap_ctx ctx, const olap_star *star, const olap_batch_query *queries,
                              int32_t nq, olap_batch *out)
{
    BatchWorkspace *ws = batch_workspace_create();
    olap_status st = batch_create_in(ctx, star, queries, nq, ws, nullptr, out);
    if (st != OLAP_OK) {
        batch_workspace_destroy(ws);
        return st;
    }
    (*out)->ws_owned = true;
    return OLAP_OK;
}

olap_status olap_batch_offsets(olap_batch b, int64_t *offsets)
{
    if (!b || !offsets) return OLAP_E_INVALID;
    std::copy(b->offsets.begin(), b->offsets.end(), offsets);
    return OLAP_OK;
}

double olap_batch_expected_events(olap_batch b) { return b ? b->events : -1.0; }

int32_t olap_batch_num_configs(void) { return kNumBatchCfgs; }

int32_t olap_batch_config_form(int32_t cfg)
{
    if (cfg < 0 || cfg >= kNumBatchCfgs) return -1;
    return kBatchCfgs[cfg].seg ? 2 : kBatchCfgs[cfg].grouped ? 1 : 0;
}

olap_status olap_batch_config_info(int32_t cfg, int32_t *threads, int32_t *rows_per_thread,
                                   int32_t *lds_table_kib, int32_t *staged)
{
    if (cfg < 0 || cfg >= kNumBatchCfgs) return OLAP_E_INVALID;
    if (threads) *threads = kBatchCfgs[cfg].bs;
    if (rows_per_thread) *rows_per_thread = kBatchCfgs[cfg].grouped ? 2 : 4 * kBatchCfgs[cfg].unroll;
    if (lds_table_kib) *lds_table_kib = kBatchCfgs[cfg].lds_kib;
    if (staged) *staged = kBatchCfgs[cfg].staged;
    return OLAP_OK;
}

olap_status olap_batch_set_config(olap_batch b, int32_t cfg)
{
    if (!b) return OLAP_E_INVALID;
    if (cfg < 0 || cfg >= kNumBatchCfgs) return bfail(b->ctx, OLAP_E_INVALID, "config out of range");
    if (!b->planned[cfg]) {
        olap_status st = plan_config(b, cfg);
        if (st != OLAP_OK) return st;
    }
    b->cfg = cfg;
    return OLAP_OK;
}

olap_status olap_batch_run(olap_batch b, int64_t *sum, int64_t *count, void *stream)
{
    if (!b) return OLAP_E_INVALID;
    if (!sum || !count) return bfail(b->ctx, OLAP_E_INVALID, "batch_run: null output");
    return batch_run(b, sum, count, (hipStream_t)stream, false);
}

olap_status olap_batch_autotune(olap_batch b, int64_t *sum, int64_t *count, void *stream,
                                int32_t *cfg_out, double *ms_out)
{
    if (!b) return OLAP_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return bfail(b->ctx, OLAP_E_HIP, "autotune events");
    int best = b->cfg;
    double best_ms = 1e30;
    olap_status st = OLAP_OK;
    for (int c = 0; c < kNumBatchCfgs && st == OLAP_OK; c++) {
        st = olap_batch_set_config(b, c);
        if (st == OLAP_E_CAPACITY) {  // this layout does not fit the batch: skip it
            st = OLAP_OK;
            continue;
        }
        if (st != OLAP_OK) break;
        st = olap_batch_run(b, sum, count, stream);
        std::vector<double> ts;
        for (int r = 0; r < 3 && st == OLAP_OK; r++) {
            (void)hipEventRecord(e0, s);
            st = olap_batch_run(b, sum, count, stream);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) st = bfail(b->ctx, OLAP_E_HIP, "autotune sync");
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ts.push_back(ms);
        }
        if (st != OLAP_OK) break;
        std::sort(ts.begin(), ts.end());
        if (ts[1] < best_ms) {
            best_ms = ts[1];
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != OLAP_OK) return st;
    b->cfg = best;
    if (cfg_out) *cfg_out = best;
    if (ms_out) *ms_out = best_ms;
    return olap_batch_run(b, sum, count, stream);
}

olap_status olap_batch_graph_capture(olap_batch b, int64_t *sum, int64_t *count, void *stream)
{
    if (!b) return OLAP_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    if (!s) return bfail(b->ctx, OLAP_E_INVALID, "graph capture needs a non-null stream");
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    // one eager run sizes the staging buffers (the record total of a batch
    // is fixed by its queries and data), then the capture skips that read
    olap_status st = batch_run(b, sum, count, s, false);
    if (st != OLAP_OK) return st;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bhip(b->ctx, e, "pre-capture run");
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bhip(b->ctx, e, "begin capture");
    st = batch_run(b, sum, count, s, true);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (st != OLAP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return bhip(b->ctx, e, "end capture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) return bhip(b->ctx, e, "graph instantiate");
    return OLAP_OK;
}

olap_status olap_batch_graph_launch(olap_batch b, void *stream)
{
    if (!b) return OLAP_E_INVALID;
    if (!b->exec) return bfail(b->ctx, OLAP_E_INVALID, "no captured graph");
    hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e != hipSuccess) return bhip(b->ctx, e, "graph launch");
    return OLAP_OK;
}

olap_status olap_batch_destroy(olap_batch b)
{
    if (!b) return OLAP_E_INVALID;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ws_owned) batch_workspace_destroy(b->ws);
    delete b;
    return OLAP_OK;
}

}  // extern "C"
