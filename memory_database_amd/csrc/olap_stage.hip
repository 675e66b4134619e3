// olap_stage.hip -- prefix and reduce passes of the staged aggregation
// (olap_stage.h), and the Stager host object that owns their scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "olap_internal.h"
#include "olap_stage.h"

namespace olap {

namespace {

constexpr int kScanThreads = 256;

// per partition (one block each): exclusive sums of its nwg counts -> starts,
// and the partition total -> totals[p]
__global__ __launch_bounds__(kScanThreads) void k_stage_scan_local(const uint32_t *counts,
                                                                   unsigned long long *starts,
                                                                   unsigned long long *totals,
                                                                   uint32_t nwg)
{
    __shared__ unsigned long long part[kScanThreads];
    const uint32_t p = blockIdx.x;
    const uint32_t per = (nwg + kScanThreads - 1) / kScanThreads;
    const uint32_t b = min(threadIdx.x * per, nwg);
    const uint32_t e = min(b + per, nwg);
    unsigned long long acc = 0;
    for (uint32_t i = b; i < e; i++) acc += counts[(size_t)p * nwg + i];
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int t = 0; t < kScanThreads; t++) {
            const unsigned long long v = part[t];
            part[t] = run;
            run += v;
        }
        totals[p] = run;
    }
    __syncthreads();
    unsigned long long run = part[threadIdx.x];
    for (uint32_t i = b; i < e; i++) {
        const unsigned long long v = counts[(size_t)p * nwg + i];
        starts[(size_t)p * nwg + i] = run;
        run += v;
    }
}

// part_begin / item_begin = exclusive sums over partitions of the record
// totals and of the reduce work items (nparts + 1 entries each)
__global__ __launch_bounds__(1024) void k_stage_scan_parts(const unsigned long long *totals,
                                                           unsigned long long *part_begin,
                                                           uint32_t *item_begin, uint32_t nparts)
{
    __shared__ unsigned long long tot[1024];
    __shared__ uint32_t items[1024];
    const uint32_t per = (nparts + 1023) / 1024;
    const uint32_t b = min(threadIdx.x * per, nparts);
    const uint32_t e = min(b + per, nparts);
    unsigned long long acc = 0;
    uint32_t it = 0;
    for (uint32_t p = b; p < e; p++) {
        acc += totals[p];
        it += (uint32_t)((totals[p] + kReduceChunk - 1) / kReduceChunk);
    }
    tot[threadIdx.x] = acc;
    items[threadIdx.x] = it;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = 0;
        uint32_t ri = 0;
        for (int t = 0; t < 1024; t++) {
            const unsigned long long v = tot[t];
            const uint32_t vi = items[t];
            tot[t] = r;
            items[t] = ri;
            r += v;
            ri += vi;
        }
        part_begin[nparts] = r;
        item_begin[nparts] = ri;
    }
    __syncthreads();
    unsigned long long r = tot[threadIdx.x];
    uint32_t ri = items[threadIdx.x];
    for (uint32_t p = b; p < e; p++) {
        part_begin[p] = r;
        item_begin[p] = ri;
        r += totals[p];
        ri += (uint32_t)((totals[p] + kReduceChunk - 1) / kReduceChunk);
    }
}

__global__ void k_stage_finalize(unsigned long long *starts, const unsigned long long *part_begin,
                                 uint32_t nparts, uint32_t nwg)
{
    const size_t n = (size_t)nparts * nwg;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        starts[i] += part_begin[i / nwg];
}

typedef __attribute__((address_space(1))) unsigned long long g_u64;

// work items: chunks of kReduceChunk records of one partition, claimed with
// one global atomic each; the partition of an item is found by binary search.
// Records are read 16 bytes per load (two 8-byte records or one 16-byte
// record), kReduceLoads independent loads per thread in flight.
constexpr int kReduceThreads = 1024;
constexpr int kReduceLoads = 8;

// claim the next work item; thread 0 finds its partition
__device__ __forceinline__ void reduce_claim(uint32_t *work, const uint32_t *item_begin, uint32_t nparts,
                                             uint32_t &s_item, uint32_t &s_part)
{
    if (threadIdx.x == 0) {
        const uint32_t item = __hip_atomic_fetch_add(work, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t lo = 0;
        if (item < item_begin[nparts]) {
            uint32_t hi = nparts;  // last p with item_begin[p] <= item
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (item_begin[mid] <= item) lo = mid; else hi = mid;
            }
        }
        s_item = item;
        s_part = lo;
    }
}

// one record into the LDS table (g >= slots: no record)
__device__ __forceinline__ void reduce_add(unsigned long long *lsum, uint32_t *lcnt, uint32_t slots, uint32_t g,
                                           long long m)
{
    if (g < slots) {
        __hip_atomic_fetch_add(&lsum[g], (unsigned long long)m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&lcnt[g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// the non-empty slots of the LDS table into the outputs. sole: this work
// item is the only one of its partition, and the scan that wrote the records
// has finished, so no one else writes these output slots now: a plain
// read-modify-write replaces the atomics
__device__ __forceinline__ void reduce_flush(const unsigned long long *lsum, const uint32_t *lcnt, uint32_t slots,
                                             unsigned long long out, unsigned long long *gsum,
                                             unsigned long long *gcnt, bool sole = false)
{
    if (sole) {
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            const uint32_t c = lcnt[s];
            if (c) {
                if (gcnt) gcnt[out + s] += (unsigned long long)c;
                gsum[out + s] += lsum[s];
            }
        }
        return;
    }
    for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
        const uint32_t c = lcnt[s];
        if (c) {
            if (gcnt)
                __hip_atomic_fetch_add((g_u64 *)&gcnt[out + s], (unsigned long long)c, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add((g_u64 *)&gsum[out + s], lsum[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Partition ranges of the two-pass form: records [part_begin[p],
// part_begin[p + 1]) of the buffer, clipped to its capacity (matches beyond
// it were added to the outputs by the scan).
template <bool WIDE>
__global__ __launch_bounds__(kReduceThreads) void k_stage_reduce(
    const unsigned long long *part_begin, const uint32_t *item_begin, uint32_t nparts,
    const void *rec, unsigned long long cap, const unsigned long long *part_out, const uint32_t *part_slots,
    uint32_t *work, unsigned long long *gsum, unsigned long long *gcnt)
{
    __shared__ unsigned long long lsum[kPartSlots];
    __shared__ uint32_t lcnt[kPartSlots];
    __shared__ uint32_t s_item, s_part;
    const uint4 *v4 = (const uint4 *)rec;
    for (;;) {
        reduce_claim(work, item_begin, nparts, s_item, s_part);
        __syncthreads();
        const uint32_t item = s_item;
        const uint32_t p = s_part;
        if (item >= item_begin[nparts]) break;
        const uint32_t slots = part_slots[p];
        const unsigned long long b = min(part_begin[p] + (unsigned long long)(item - item_begin[p]) * kReduceChunk, cap);
        const unsigned long long e = min(min(b + kReduceChunk, part_begin[p + 1]), cap);
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            lsum[s] = 0ull;
            lcnt[s] = 0u;
        }
        __syncthreads();
        // 16-byte units [u0, u1): record pairs from the even record at or
        // below b (narrow), or single records (wide); cap is even, so the
        // pair holding record e - 1 ends within the buffer
        const unsigned long long u0 = WIDE ? b : (b >> 1);
        const unsigned long long u1 = WIDE ? e : ((e + 1) >> 1);
        constexpr unsigned long long step = (unsigned long long)kReduceThreads * kReduceLoads;
        for (unsigned long long i0 = u0 + threadIdx.x; i0 < u1; i0 += step) {
            uint4 v[kReduceLoads];
#pragma unroll
            for (int u = 0; u < kReduceLoads; u++) {
                const unsigned long long i = i0 + (unsigned long long)u * kReduceThreads;
                v[u] = i < u1 ? v4[i] : make_uint4(kPartSlots, 0u, kPartSlots, 0u);
            }
#pragma unroll
            for (int u = 0; u < kReduceLoads; u++) {
                const unsigned long long i = i0 + (unsigned long long)u * kReduceThreads;
                if (WIDE) {
                    reduce_add(lsum, lcnt, slots, v[u].x, (long long)(int32_t)v[u].y + (long long)(int32_t)v[u].z);
                } else {
                    const unsigned long long r = 2ull * i;  // record index of v[u].x
                    reduce_add(lsum, lcnt, slots, r >= b && r < e ? v[u].x : kPartSlots, (long long)(int32_t)v[u].y);
                    reduce_add(lsum, lcnt, slots, r + 1 >= b && r + 1 < e ? v[u].z : kPartSlots,
                               (long long)(int32_t)v[u].w);
                }
            }
        }
        __syncthreads();
        reduce_flush(lsum, lcnt, slots, part_out[p], gsum, gcnt, item_begin[p + 1] - item_begin[p] == 1);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- one-pass
// chunks handed out by a scan: ids below min(cursor, cap) / chunk size

__device__ __forceinline__ uint32_t chunks_used(const unsigned long long *cursor, unsigned long long cap,
                                                uint32_t ch_bits)
{
    return (uint32_t)(min(*cursor, cap) >> ch_bits);
}

// chunk counts per partition: a per-workgroup LDS histogram (partitions
// beyond it count with global atomics directly), then one atomic per bin
constexpr uint32_t kHistBins = 8192;

__global__ __launch_bounds__(256) void k_chunk_count(const unsigned long long *cursor, unsigned long long cap,
                                                     uint32_t ch_bits, const uint32_t *chunk_part,
                                                     uint32_t *part_chunks, uint32_t nparts)
{
    __shared__ uint32_t hist[kHistBins];
    const uint32_t nb = min(nparts, kHistBins);
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0u;
    __syncthreads();
    const uint32_t n = chunks_used(cursor, cap, ch_bits);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const uint32_t p = chunk_part[c];
        if (p < nb) atomicAdd(&hist[p], 1u);
        else atomicAdd(&part_chunks[p], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        if (hist[i]) atomicAdd(&part_chunks[i], hist[i]);
}

// exclusive starts of every partition's chunks (in place, total at
// [nparts]), scatter cursors, and reduce work items of cpi chunks
__global__ __launch_bounds__(1024) void k_chunk_scan(uint32_t *part_chunks, uint32_t *part_ccur,
                                                     uint32_t *item_begin, uint32_t nparts, uint32_t cpi)
{
    __shared__ uint32_t tot[1024];
    __shared__ uint32_t items[1024];
    const uint32_t per = (nparts + 1023) / 1024;
    const uint32_t b = min(threadIdx.x * per, nparts);
    const uint32_t e = min(b + per, nparts);
    uint32_t acc = 0, it = 0;
    for (uint32_t p = b; p < e; p++) {
        acc += part_chunks[p];
        it += (part_chunks[p] + cpi - 1) / cpi;
    }
    tot[threadIdx.x] = acc;
    items[threadIdx.x] = it;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0, ri = 0;
        for (int t = 0; t < 1024; t++) {
            const uint32_t v = tot[t], vi = items[t];
            tot[t] = r;
            items[t] = ri;
            r += v;
            ri += vi;
        }
        part_chunks[nparts] = r;
        item_begin[nparts] = ri;
    }
    __syncthreads();
    uint32_t r = tot[threadIdx.x], ri = items[threadIdx.x];
    for (uint32_t p = b; p < e; p++) {
        const uint32_t c = part_chunks[p];
        part_chunks[p] = r;
        part_ccur[p] = r;
        item_begin[p] = ri;
        r += c;
        ri += (c + cpi - 1) / cpi;
    }
}

// chunk ids grouped by partition, each with its fill: (fill << 32) | id;
// the same per-workgroup LDS pre-count reserves each bin's slots at once
__global__ __launch_bounds__(256) void k_chunk_scatter(const unsigned long long *cursor, unsigned long long cap,
                                                       uint32_t ch_bits, const uint32_t *chunk_part,
                                                       const uint32_t *chunk_fill, uint32_t *part_ccur,
                                                       unsigned long long *order, uint32_t nparts)
{
    __shared__ uint32_t hist[kHistBins];
    const uint32_t nb = min(nparts, kHistBins);
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0u;
    __syncthreads();
    const uint32_t n = chunks_used(cursor, cap, ch_bits);
    const uint32_t stride = gridDim.x * blockDim.x;
    // pass 1: count this workgroup's chunks per bin
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
        const uint32_t p = chunk_part[c];
        if (p < nb) atomicAdd(&hist[p], 1u);
    }
    __syncthreads();
    // reserve each bin's slots with one global atomic
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        if (hist[i]) hist[i] = atomicAdd(&part_ccur[i], hist[i]);
    __syncthreads();
    // pass 2: place the chunks (LDS cursors inside the reserved ranges)
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
        const uint32_t p = chunk_part[c];
        const uint32_t pos = p < nb ? atomicAdd(&hist[p], 1u) : atomicAdd(&part_ccur[p], 1u);
        order[pos] = ((unsigned long long)chunk_fill[c] << 32) | c;
    }
}

// A work item's chunks laid end to end by their filled 16-byte units (kper
// records a unit): ustart[c] = units before chunk c, ustart[n] the total.
// Wave 0 scans; the caller synchronises before use. A partly filled chunk
// (a workgroup's last chunk of a partition) then costs only its records.
__device__ __forceinline__ void unit_starts(const unsigned long long *lord, uint32_t n, uint32_t kper,
                                            uint32_t *ustart)
{
    if (threadIdx.x >= 64u) return;
    const uint32_t lane = threadIdx.x;
    uint32_t carry = 0u;
    for (uint32_t b = 0; b < n; b += 64u) {
        const uint32_t c = b + lane;
        uint32_t v = c < n ? ((uint32_t)(lord[c] >> 32) + kper - 1u) / kper : 0u;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(v, (unsigned)d);
            if (lane >= (uint32_t)d) v += t;
        }
        if (c < n) ustart[c + 1u] = carry + v;
        carry += __shfl(v, 63);
    }
    if (lane == 0u) ustart[0] = 0u;
}

// the chunk of the item holding filled unit i (ustart[c] <= i < ustart[c+1]),
// by a branch-free binary search over ustart[0..n]
__device__ __forceinline__ uint32_t unit_chunk(const uint32_t *ustart, uint32_t n, uint32_t i)
{
    uint32_t c = 0u;
    for (uint32_t step = 1u << (31 - __clz(n)); step; step >>= 1)
        if (c + step <= n && ustart[c + step] <= i) c += step;
    return c;
}

// work items: up to cpi chunks of one partition, aggregated in an LDS table
template <uint32_t FMT>
__device__ __forceinline__ void chunk_reduce(
    const uint32_t *part_chunks, const uint32_t *item_begin, uint32_t nparts, uint32_t cpi,
    const unsigned long long *order, const void *pool, uint32_t ch_bits, int32_t mbase,
    const unsigned long long *part_out, const uint32_t *part_slots, uint32_t *work,
    unsigned long long *gsum, unsigned long long *gcnt)
{
    __shared__ unsigned long long lsum[kPartSlots];
    __shared__ uint32_t lcnt[kPartSlots];
    __shared__ unsigned long long lord[kReduceItemChunks];  // the item's (fill << 32 | chunk id)
    __shared__ uint32_t ustart[kReduceItemChunks + 1];
    __shared__ uint32_t s_item, s_part;
    const uint4 *v4 = (const uint4 *)pool;
    // 16-byte units of a chunk: one WIDE record; four COMPACT records; or
    // four NARROW records, as 16 bytes of measures and 8 bytes of slots
    constexpr uint32_t kPer = FMT == REC_WIDE ? 1u : 4u;
    const unsigned long long ch = 1ull << ch_bits;
    for (;;) {
        reduce_claim(work, item_begin, nparts, s_item, s_part);
        __syncthreads();
        const uint32_t item = s_item;
        const uint32_t p = s_part;
        if (item >= item_begin[nparts]) break;
        const uint32_t slots = part_slots[p];
        const uint32_t cb = part_chunks[p] + (item - item_begin[p]) * cpi;
        const uint32_t n = min(cb + cpi, part_chunks[p + 1]) - cb;  // at most cpi <= kReduceItemChunks
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lord[i] = order[cb + i];
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            lsum[s] = 0ull;
            lcnt[s] = 0u;
        }
        __syncthreads();
        unit_starts(lord, n, kPer, ustart);
        __syncthreads();
        // items of mostly full chunks walk every unit (a bounds check skips
        // the rest of the last ones); sparse items only the filled units
        const uint32_t ubits = FMT == REC_WIDE ? ch_bits : ch_bits - 2u;
        const bool filled_only = 2u * ustart[n] < (n << ubits);
        const uint32_t units = filled_only ? ustart[n] : n << ubits;
        for (uint32_t i0 = threadIdx.x; i0 < units; i0 += (uint32_t)kReduceThreads * kReduceLoads) {
            uint4 v[kReduceLoads];
            uint2 gs[kReduceLoads];
            uint32_t r[kReduceLoads], fill[kReduceLoads];
#pragma unroll
            for (int u = 0; u < kReduceLoads; u++) {
                const uint32_t i = i0 + (uint32_t)u * kReduceThreads;
                v[u] = make_uint4(0u, 0u, 0u, 0u);
                gs[u] = make_uint2(0u, 0u);
                r[u] = 0u;
                fill[u] = 0u;
                if (i < units) {
                    const uint32_t c = filled_only ? unit_chunk(ustart, n, i) : i >> ubits;
                    const unsigned long long o = lord[c];
                    const uint32_t w = filled_only ? i - ustart[c] : i & ((1u << ubits) - 1u);
                    const unsigned long long cbase = (unsigned long long)(uint32_t)o << ch_bits;  // first record
                    r[u] = kPer * w;  // first record index of the unit within the chunk
                    fill[u] = (uint32_t)(o >> 32);
                    if (r[u] < fill[u]) {
                        if (FMT == REC_WIDE) {
                            v[u] = v4[cbase + w];
                        } else if (FMT == REC_COMPACT) {
                            v[u] = v4[(cbase >> 2) + w];
                        } else {
                            const char *cp = (const char *)pool + 6ull * cbase;
                            v[u] = ((const uint4 *)cp)[w];
                            gs[u] = ((const uint2 *)(cp + 4ull * ch))[w];
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kReduceLoads; u++) {
                if (FMT == REC_WIDE) {
                    reduce_add(lsum, lcnt, slots, r[u] < fill[u] ? v[u].x : kPartSlots,
                               (long long)(int32_t)v[u].y + (long long)(int32_t)v[u].z);
                } else {
                    const uint32_t x[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        uint32_t g;
                        long long m;
                        if (FMT == REC_COMPACT) {
                            g = x[k] >> kCompactMeasureBits;
                            m = (long long)mbase + (long long)(x[k] & ((1u << kCompactMeasureBits) - 1u));
                        } else {
                            g = k == 0 ? (gs[u].x & 0xffffu) : k == 1 ? (gs[u].x >> 16)
                                : k == 2 ? (gs[u].y & 0xffffu) : (gs[u].y >> 16);
                            m = (long long)(int32_t)x[k];
                        }
                        reduce_add(lsum, lcnt, slots, r[u] + (uint32_t)k < fill[u] ? g : kPartSlots, m);
                    }
                }
            }
        }
        __syncthreads();
        reduce_flush(lsum, lcnt, slots, part_out[p], gsum, gcnt);
        __syncthreads();
    }
}

#define CHUNK_REDUCE_PARAMS                                                                              \
    const uint32_t *part_chunks, const uint32_t *item_begin, uint32_t nparts, uint32_t cpi,             \
        const unsigned long long *order, const void *pool, uint32_t ch_bits, int32_t mbase,             \
        const unsigned long long *part_out, const uint32_t *part_slots, uint32_t *work,                 \
        unsigned long long *gsum, unsigned long long *gcnt
#define CHUNK_REDUCE_ARGS \
    part_chunks, item_begin, nparts, cpi, order, pool, ch_bits, mbase, part_out, part_slots, work, gsum, gcnt

template <uint32_t FMT>
__global__ __launch_bounds__(kReduceThreads) void k_chunk_reduce(CHUNK_REDUCE_PARAMS)
{
    chunk_reduce<FMT>(CHUNK_REDUCE_ARGS);
}

// 4-byte records (slot << 20 | measure - mbase): one 64-bit LDS atomic a
// record adds (measure offset << 22) | 1, so a slot's word holds its count
// in the low 22 bits and its sum of offsets above them. A work item holds
// at most kReduceChunk = 2^21 records, so neither field overflows (count
// <= 2^21 < 2^22, sum < 2^21 * 2^20 = 2^41 < 2^42); the flush adds
// count * mbase back. A table of 2^12 slots (32 KiB) and at most 64 VGPRs
// let two 1024-thread workgroups share a CU.
constexpr uint32_t kPackCountBits = 22;
static_assert(kReduceChunk < (1u << kPackCountBits), "packed count field");
static_assert(kCompactMeasureBits + 21 + kPackCountBits <= 64, "packed sum field");

__global__ __launch_bounds__(kReduceThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_chunk_reduce_compact(
    CHUNK_REDUCE_PARAMS)
{
    constexpr uint32_t kSlots = 1u << kCompactPartBits;
    __shared__ unsigned long long lpack[kSlots];
    __shared__ unsigned long long lord[kReduceItemChunks];  // the item's (fill << 32 | chunk id)
    __shared__ uint32_t s_item, s_part;
    typedef __attribute__((address_space(3))) unsigned long long l_u64;
    const uint4 *v4 = (const uint4 *)pool;
    for (;;) {
        reduce_claim(work, item_begin, nparts, s_item, s_part);
        __syncthreads();
        const uint32_t item = s_item;
        const uint32_t p = s_part;
        if (item >= item_begin[nparts]) break;
        const uint32_t slots = part_slots[p];
        const uint32_t cb = part_chunks[p] + (item - item_begin[p]) * cpi;
        const uint32_t n = min(cb + cpi, part_chunks[p + 1]) - cb;  // at most cpi <= kReduceItemChunks
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lord[i] = order[cb + i];
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) lpack[s] = 0ull;
        __syncthreads();
        // (every unit of the item's chunks: the batched scan's chunks are
        // mostly full, and chunk_reduce's walk over filled units only would
        // spill registers here)
        const uint32_t ubits = ch_bits - 2u;  // 16-byte units of a chunk: four records
        const uint32_t units = n << ubits;
        for (uint32_t i0 = threadIdx.x; i0 < units; i0 += (uint32_t)kReduceThreads * kReduceLoads) {
            uint4 v[kReduceLoads];
            uint32_t r[kReduceLoads], fill[kReduceLoads];
#pragma unroll
            for (int u = 0; u < kReduceLoads; u++) {
                const uint32_t i = i0 + (uint32_t)u * kReduceThreads;
                v[u] = make_uint4(0u, 0u, 0u, 0u);
                r[u] = 0u;
                fill[u] = 0u;
                if (i < units) {
                    const unsigned long long o = lord[i >> ubits];
                    const uint32_t w = i & ((1u << ubits) - 1u);
                    r[u] = 4u * w;
                    fill[u] = (uint32_t)(o >> 32);
                    if (r[u] < fill[u]) v[u] = v4[(((unsigned long long)(uint32_t)o << ch_bits) >> 2) + w];
                }
            }
#pragma unroll
            for (int u = 0; u < kReduceLoads; u++) {
                const uint32_t x[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t g = x[k] >> kCompactMeasureBits;
                    if (r[u] + (uint32_t)k < fill[u] && g < slots) {
                        const unsigned long long add =
                            ((unsigned long long)(x[k] & ((1u << kCompactMeasureBits) - 1u)) << kPackCountBits) | 1ull;
                        __hip_atomic_fetch_add((l_u64 *)&lpack[g], add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
            }
        }
        __syncthreads();
        const unsigned long long out = part_out[p];
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            const unsigned long long w = lpack[s];
            const unsigned long long c = w & ((1ull << kPackCountBits) - 1ull);
            if (c) {
                const unsigned long long sum = (w >> kPackCountBits) + c * (unsigned long long)(long long)mbase;
                if (gcnt)
                    __hip_atomic_fetch_add((g_u64 *)&gcnt[out + s], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add((g_u64 *)&gsum[out + s], sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
    }
}

// grow-only device array: reallocated only when n exceeds its capacity
template <typename T>
hipError_t grow_dev(T *&p, size_t &cap, size_t n)
{
    n = std::max<size_t>(n, 1);
    if (p && n <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e == hipSuccess) cap = n;
    return e;
}

}  // namespace

hipError_t Stager::setup(const unsigned long long *out_off, const uint32_t *G, int nq, uint32_t nwg_,
                         bool two, hipStream_t s, uint32_t pbits)
{
    qbase_.assign((size_t)nq, 0);
    std::vector<unsigned long long> pout;
    std::vector<uint32_t> pslots;
    const uint32_t ps = 1u << pbits;
    for (int q = 0; q < nq; q++) {
        qbase_[(size_t)q] = (uint32_t)pout.size();
        const uint32_t np = (G[q] + ps - 1) / ps;
        for (uint32_t i = 0; i < np; i++) {
            pout.push_back(out_off[q] + (unsigned long long)i * ps);
            pslots.push_back(std::min<uint32_t>(ps, G[q] - i * ps));
        }
    }
    part_bits = pbits;
    if (pout.size() > kMaxParts) return hipErrorInvalidValue;
    nparts = (uint32_t)pout.size();
    nwg = nwg_;
    two_measures = two;
    const size_t np = std::max<uint32_t>(nparts, 1);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = grow_dev(wg_counts, cap_[0], np * nwg);
    if (e == hipSuccess) e = grow_dev(wg_start, cap_[1], np * nwg);
    if (e == hipSuccess) e = grow_dev(part_tot, cap_[2], np);
    if (e == hipSuccess) e = grow_dev(part_begin, cap_[3], np + 1);
    if (e == hipSuccess) e = grow_dev(item_begin, cap_[4], np + 1);
    if (e == hipSuccess) e = grow_dev(work, cap_[5], 4);
    if (e == hipSuccess) e = grow_dev(part_out, cap_[6], np);
    if (e == hipSuccess) e = grow_dev(part_slots, cap_[7], np);
    if (e == hipSuccess && nparts) {
        e = hipMemcpyAsync(part_out, pout.data(), nparts * sizeof(unsigned long long), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(part_slots, pslots.data(), nparts * sizeof(uint32_t), hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess) e = launch_zero_u32(work, 4, s);
    // the host vectors are read by the copies: wait for them
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

hipError_t Stager::prefix(hipStream_t s)
{
    if (!nparts) return hipSuccess;
    k_stage_scan_local<<<nparts, kScanThreads, 0, s>>>(wg_counts, wg_start, part_tot, nwg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_stage_scan_parts<<<1, 1024, 0, s>>>(part_tot, part_begin, item_begin, nparts);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t n = (size_t)nparts * nwg;
    k_stage_finalize<<<(unsigned)std::min<size_t>((n + 255) / 256, 4096), 256, 0, s>>>(
        wg_start, part_begin, nparts, nwg);
    return hipGetLastError();
}

hipError_t Stager::total(unsigned long long *out, hipStream_t s)
{
    *out = 0;
    if (!nparts) return hipSuccess;
    hipError_t e = hipMemcpyAsync(out, part_begin + nparts, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

hipError_t Stager::reserve(unsigned long long records)
{
    if (records <= cap && rec && rec_wide == two_measures) return hipSuccess;
    // even, so the reduce's 16-byte record pairs never read past it
    const unsigned long long want = (records + records / 4 + 4096) & ~1ull;
    if (rec) (void)hipFree(rec);
    rec = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&rec, want * (two_measures ? 16u : 8u));
    if (e == hipSuccess) {
        cap = want;
        rec_wide = two_measures;
    }
    return e;
}

hipError_t Stager::reduce(unsigned long long *gsum, unsigned long long *gcnt, int num_cus,
                          hipStream_t s)
{
    if (!nparts) return hipSuccess;
    hipError_t e = launch_zero_u32(work, 1, s);
    if (e != hipSuccess) return e;
    // 96 KB of LDS per workgroup: one per CU
    if (rec_wide)
        k_stage_reduce<true><<<num_cus, kReduceThreads, 0, s>>>(part_begin, item_begin, nparts, rec, cap,
                                                                part_out, part_slots, work, gsum, gcnt);
    else
        k_stage_reduce<false><<<num_cus, kReduceThreads, 0, s>>>(part_begin, item_begin, nparts, rec, cap,
                                                                 part_out, part_slots, work, gsum, gcnt);
    return hipGetLastError();
}

hipError_t Stager::chunk_reserve(unsigned long long records, uint32_t nwg_, uint32_t ch_bits, uint32_t fmt,
                                 int32_t mbase)
{
    const unsigned long long ch = 1ull << ch_bits;
    // expected matches + 25%, plus the two chunks every workgroup holds open
    // in every partition (a partial current chunk and a spare)
    unsigned long long want =
        records + records / 4 + 2ull * (unsigned long long)std::max<uint32_t>(nparts, 1) * nwg_ * ch + ch;
    // a pool that has to grow grows by at least half, so a workload of
    // growing batches reallocates a few times rather than at every batch
    if (pool && want > pool_cap) want = std::max<unsigned long long>(want, pool_cap + pool_cap / 2);
    want = (want + ch - 1) & ~(ch - 1);
    hipError_t e = hipSuccess;
    if (!pool_cursor) e = hipMalloc(&pool_cursor, sizeof(unsigned long long));
    static const uint32_t rec_bytes[3] = {6u, 16u, 4u};  // NARROW, WIDE, COMPACT
    if (e == hipSuccess &&
        (want > pool_cap || pool_ch_bits != ch_bits || rec_bytes[pool_fmt] < rec_bytes[fmt] || !pool)) {
        if (pool) (void)hipFree(pool);
        if (chunk_part) (void)hipFree(chunk_part);
        if (chunk_fill) (void)hipFree(chunk_fill);
        if (chunk_order) (void)hipFree(chunk_order);
        pool = nullptr;
        chunk_part = chunk_fill = nullptr;
        chunk_order = nullptr;
        pool_cap = 0;
        const size_t nch = (size_t)(want >> ch_bits);
        e = hipMalloc(&pool, (size_t)want * rec_bytes[fmt]);
        if (e == hipSuccess) e = hipMalloc(&chunk_part, nch * 4);
        if (e == hipSuccess) e = hipMalloc(&chunk_fill, nch * 4);
        if (e == hipSuccess) e = hipMalloc(&chunk_order, nch * 8);
        if (e == hipSuccess) {
            pool_cap = want;
            pool_ch_bits = ch_bits;
        }
    }
    if (e == hipSuccess) {
        // a pool allocated for a wider format holds a narrower one
        pool_fmt = fmt;
        pool_mbase = mbase;
    }
    const size_t np = std::max<uint32_t>(nparts, 1);
    if (e == hipSuccess) e = grow_dev(part_chunks, ccap_[0], np + 1);
    if (e == hipSuccess) e = grow_dev(part_ccur, ccap_[1], np);
    if (e == hipSuccess) e = grow_dev(citem_begin, ccap_[2], np + 1);
    return e;
}

hipError_t Stager::chunk_clear(hipStream_t s)
{
    hipError_t e = launch_zero_u64(pool_cursor, 1, s);
    if (e == hipSuccess) e = launch_zero_u32(work, 4, s);
    return e;
}

ChunkArgs Stager::chunk_args() const
{
    ChunkArgs c;
    c.pool = pool;
    c.cap = pool_cap;
    c.cursor = pool_cursor;
    c.chunk_part = chunk_part;
    c.chunk_fill = chunk_fill;
    c.part_out = part_out;
    c.overflow = work + 1;
    c.ch_bits = pool_ch_bits;
    c.nparts = nparts;
    c.fmt = pool_fmt;
    c.mbase = pool_mbase;
    return c;
}

hipError_t Stager::chunk_reduce(unsigned long long *gsum, unsigned long long *gcnt, int num_cus, hipStream_t s)
{
    if (!nparts) return hipSuccess;
    const unsigned long long nch = pool_cap >> pool_ch_bits;
    const unsigned grid = (unsigned)std::max<unsigned long long>(1, std::min<unsigned long long>((nch + 255) / 256, 4096));
    hipError_t e = launch_zero_u32(part_chunks, (size_t)nparts + 1, s);
    if (e == hipSuccess) {
        k_chunk_count<<<grid, 256, 0, s>>>(pool_cursor, pool_cap, pool_ch_bits, chunk_part, part_chunks, nparts);
        e = hipGetLastError();
    }
    // reduce items of about kReduceChunk records
    const uint32_t cpi = std::min<uint32_t>(kReduceItemChunks, std::max<uint32_t>(1u, (uint32_t)(kReduceChunk >> pool_ch_bits)));
    if (e == hipSuccess) {
        k_chunk_scan<<<1, 1024, 0, s>>>(part_chunks, part_ccur, citem_begin, nparts, cpi);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        k_chunk_scatter<<<grid, 256, 0, s>>>(pool_cursor, pool_cap, pool_ch_bits, chunk_part, chunk_fill, part_ccur,
                                             chunk_order, nparts);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = launch_zero_u32(work, 1, s);
    if (e != hipSuccess) return e;
    auto launch = [&](auto kernel) {
        const int per_cu = pool_fmt == REC_COMPACT ? 2 : 1;
        kernel<<<num_cus * per_cu, kReduceThreads, 0, s>>>(part_chunks, citem_begin, nparts, cpi, chunk_order,
                                                           pool, pool_ch_bits, pool_mbase, part_out, part_slots,
                                                           work, gsum, gcnt);
    };
    if (pool_fmt == REC_WIDE) launch(k_chunk_reduce<REC_WIDE>);
    else if (pool_fmt == REC_COMPACT) launch(k_chunk_reduce_compact);
    else launch(k_chunk_reduce<REC_NARROW>);
    return hipGetLastError();
}

StageArgs Stager::args(uint32_t cur_lds, uint32_t start_lds, unsigned long long *gsum,
                        unsigned long long *gcnt) const
{
    StageArgs a;
    a.nparts = nparts;
    a.nwg = nwg;
    a.cur_lds = cur_lds;
    a.start_lds = start_lds;
    a.wg_counts = wg_counts;
    a.wg_start = wg_start;
    a.rec2 = rec_wide ? nullptr : (uint2 *)rec;
    a.rec4 = rec_wide ? (uint4 *)rec : nullptr;
    a.cap = cap;
    a.overflow = work + 1;
    a.part_out = part_out;
    a.gsum = gsum;
    a.gcnt = gcnt;
    return a;
}

uint32_t Stager::lds_bytes() const
{
    // u32 counters/cursors, then 8-byte aligned u64 range starts
    return ((nparts * 4u + 15u) & ~15u) + nparts * 8u;
}

hipError_t Stager::overflowed(uint32_t *out, hipStream_t s)
{
    *out = 0;
    if (!work) return hipSuccess;
    hipError_t e = hipMemcpyAsync(out, work + 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

void Stager::release()
{
    void *ptrs[] = {wg_counts, wg_start, part_tot, part_begin, item_begin,
                    work,      part_out, part_slots, rec};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    wg_counts = nullptr;
    wg_start = nullptr;
    part_tot = nullptr;
    part_begin = nullptr;
    item_begin = nullptr;
    work = nullptr;
    part_out = nullptr;
    part_slots = nullptr;
    rec = nullptr;
    cap = 0;
    nparts = 0;
    for (size_t &c : cap_) c = 0;
    void *cptrs[] = {pool, pool_cursor, chunk_part, chunk_fill, chunk_order, part_chunks, part_ccur, citem_begin};
    for (void *p : cptrs)
        if (p) (void)hipFree(p);
    pool = nullptr;
    pool_cursor = nullptr;
    chunk_part = chunk_fill = part_chunks = part_ccur = citem_begin = nullptr;
    chunk_order = nullptr;
    pool_cap = 0;
    pool_ch_bits = 0;
    for (size_t &c : ccap_) c = 0;
}

}  // namespace olap
