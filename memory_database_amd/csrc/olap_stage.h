// olap_stage.h -- staged (radix-partitioned) group-by aggregation.
//
// Global 64-bit atomics are far slower than the LDS atomics of a
// workgroup's table, so group tables that do not fit in LDS are aggregated
// in three passes instead of with one global atomic pair per matching row:
//   COUNT   the scan counts its matches per (workgroup, partition) in LDS;
//           a partition is kPartSlots consecutive gids of one query;
//   prefix  exclusive sums give every (partition, workgroup) a private range
//           of the record buffer and every partition its record range;
//   SCATTER the same scan (same grid, same tile order) writes one record per
//           match -- gid within the partition and the raw measure(s) -- to
//           its workgroup's range, with LDS cursors (no global atomics);
//   reduce  workgroups take chunks of a partition's records, aggregate them
//           in an LDS table of kPartSlots slots and add the non-empty slots
//           to the int64 outputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace olap {

constexpr uint32_t kPartBits = 13;
constexpr uint32_t kPartSlots = 1u << kPartBits;       // 8192 gids per partition (96 KB LDS table)
constexpr uint32_t kReduceChunk = 1u << 21;             // records per reduce work item
constexpr uint32_t kReduceItemChunks = 2048;            // one-pass: at most this many chunks per work item
constexpr uint32_t kMaxParts = 1u << 16;

enum AggMode : int { AGG_DIRECT = 0, AGG_COUNT = 1, AGG_SCATTER = 2, AGG_CHUNK = 3 };

// One-pass record formats: a chunk of CH records holds CH u32 measures then
// CH u16 slots (NARROW, one measure, 6 bytes a record), CH uint4 {slot, m1,
// m2, 0} (WIDE, two measures), or CH u32 slot << 20 | (measure - base)
// (COMPACT: one measure whose range spans fewer than 2^20 values, with
// partitions of 2^12 slots; 4 bytes a record)
enum RecFmt : uint32_t { REC_NARROW = 0, REC_WIDE = 1, REC_COMPACT = 2 };
constexpr uint32_t kCompactPartBits = 12;
constexpr uint32_t kCompactMeasureBits = 20;
constexpr uint32_t kNoChunk = 0xffffffffu;   // no chunk (the pool is used up)

// One-pass staging (AGG_CHUNK): instead of a COUNT pass that sizes exact
// ranges, a workgroup keeps per partition a current chunk of a shared pool
// and a spare one, and hands out record slots with an LDS counter: positions
// past the current chunk's end fall into the spare, so every match writes its
// record at once. Between rounds of matches, full chunks retire, the spare
// becomes current and a new spare is claimed with one global atomic. Each
// chunk records its partition and fill; after the scan the chunks are grouped
// by partition and reduced like the ranges of the two-pass form. A pool that
// runs out (it is sized from the expected matches) never loses a match: the
// scan adds such matches to the outputs with global atomics.
struct ChunkArgs {
    void *pool;                       // chunks of CH records in format fmt
    unsigned long long cap;           // records the pool holds
    unsigned long long *cursor;       // records handed out (whole chunks)
    uint32_t *chunk_part;             // [cap / chunk] partition of each chunk
    uint32_t *chunk_fill;             // [cap / chunk] records written
    const unsigned long long *part_out;  // first output slot of each partition
    uint32_t *overflow;               // matches that went to global atomics
    uint32_t ch_bits;                 // log2 of the chunk size
    uint32_t nparts;
    uint32_t fmt;                     // RecFmt
    int32_t mbase;                    // COMPACT: the measure of offset 0
};

struct StageArgs {
    uint32_t nparts;                  // partitions of this scan
    uint32_t nwg;                     // workgroups of the scan (both passes)
    uint32_t cur_lds;                 // word offset: u32 counter / cursor per partition
    uint32_t start_lds;               // word offset: u64 range start per partition (scatter)
    uint32_t *wg_counts;              // [nparts][nwg]
    const unsigned long long *wg_start;  // [nparts][nwg]
    // one record per match, written with a single store: {gid within the
    // partition, m1} (uint2), or {gid, m1, m2, 0} (uint4) with two measures
    uint2 *rec2;
    uint4 *rec4;
    unsigned long long cap;           // records the buffer holds
    uint32_t *overflow;               // matches beyond cap (added with global atomics)
    // outputs, for matches beyond cap (a buffer sized by an earlier scan of
    // columns that have since changed): never lost, added directly
    const unsigned long long *part_out;  // first output slot of each partition
    unsigned long long *gsum;
    unsigned long long *gcnt;         // null: sums only
};

#if defined(__HIPCC__)
// LDS accesses with per-lane indexes go through explicit address-space-3
// pointers (generic pointers derived from LDS trip a gfx950 code-generation
// error in this compiler inside the batched kernel's event loop)
typedef __attribute__((address_space(3))) uint32_t stage_l_u32;
typedef __attribute__((address_space(3))) unsigned long long stage_l_u64;

// count one match of partition p (LDS counter)
__device__ __forceinline__ void stage_count(const StageArgs &s, uint32_t *lds, uint32_t p)
{
    stage_l_u32 *c = (stage_l_u32 *)lds + s.cur_lds;
    __hip_atomic_fetch_add(c + p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// a match with no room in the record buffer goes to the outputs directly
__device__ __forceinline__ void stage_overflow(const StageArgs &s, uint32_t p, uint32_t gid_low, long long m)
{
    typedef __attribute__((address_space(1))) unsigned long long s_g_u64;
    const unsigned long long o = s.part_out[p] + gid_low;
    __hip_atomic_fetch_add((s_g_u64 *)&s.gsum[o], (unsigned long long)m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s.gcnt) __hip_atomic_fetch_add((s_g_u64 *)&s.gcnt[o], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(s.overflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one-pass staging: claim a chunk of the pool for partition p; its id (first
// record id >> ch_bits), or kNoChunk when the pool is used up
__device__ __forceinline__ uint32_t chunk_claim(const ChunkArgs &c, uint32_t p)
{
    const unsigned long long ch = 1ull << c.ch_bits;
    const unsigned long long nc =
        __hip_atomic_fetch_add(c.cursor, ch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (nc + ch > c.cap) return kNoChunk;
    const uint32_t id = (uint32_t)(nc >> c.ch_bits);
    c.chunk_part[id] = p;
    return id;
}

// write one record of partition p
__device__ __forceinline__ void stage_scatter(const StageArgs &s, uint32_t *lds, uint32_t p,
                                              uint32_t gid_low, int32_t m1, int32_t m2)
{
    stage_l_u32 *c = (stage_l_u32 *)lds + s.cur_lds;
    const uint32_t pos = __hip_atomic_fetch_add(c + p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const stage_l_u64 *st = (const stage_l_u64 *)((stage_l_u32 *)lds + s.start_lds);
    const unsigned long long i = st[p] + pos;
    if (i >= s.cap) {
        stage_overflow(s, p, gid_low, (long long)m1 + (long long)m2);
        return;
    }
    if (s.rec4) s.rec4[i] = make_uint4(gid_low, (uint32_t)m1, (uint32_t)m2, 0u);
    else s.rec2[i] = make_uint2(gid_low, (uint32_t)m1);
}

// pass prologue: zero the counters (count) or load the range starts (scatter)
template <int MODE>
__device__ __forceinline__ void stage_begin(const StageArgs &s, uint32_t *lds)
{
    if (MODE == AGG_DIRECT) return;
    for (uint32_t p = threadIdx.x; p < s.nparts; p += blockDim.x) {
        lds[s.cur_lds + p] = 0u;
        if (MODE == AGG_SCATTER) {
            unsigned long long *st = (unsigned long long *)(lds + s.start_lds);
            st[p] = s.wg_start[(size_t)p * s.nwg + blockIdx.x];
        }
    }
}

// pass epilogue of COUNT: publish this workgroup's counts (after a barrier)
template <int MODE>
__device__ __forceinline__ void stage_end(const StageArgs &s, uint32_t *lds)
{
    if (MODE != AGG_COUNT) return;
    for (uint32_t p = threadIdx.x; p < s.nparts; p += blockDim.x)
        s.wg_counts[(size_t)p * s.nwg + blockIdx.x] = lds[s.cur_lds + p];
}
#endif

// Host side, olap_stage.hip. A Stager owns the scratch of one scan.
struct Stager {
    uint32_t nparts = 0;
    uint32_t nwg = 0;
    uint32_t part_bits = kPartBits;
    bool two_measures = false;
    uint32_t *wg_counts = nullptr;              // [nparts][nwg]
    unsigned long long *wg_start = nullptr;     // [nparts][nwg]
    unsigned long long *part_tot = nullptr;     // records per partition
    unsigned long long *part_begin = nullptr;   // nparts + 1
    uint32_t *item_begin = nullptr;             // nparts + 1
    uint32_t *work = nullptr;                   // [0] reduce work counter, [1] overflow
    unsigned long long *part_out = nullptr;     // first output slot of each partition
    uint32_t *part_slots = nullptr;             // slots of each partition
    void *rec = nullptr;                        // uint2 or uint4 records
    bool rec_wide = false;                      // rec holds uint4 records
    unsigned long long cap = 0;

    // partitions of every query: out_off[q] (first output slot) and G[q];
    // buffers only grow, so a Stager reused across scans stops allocating;
    // the uploads and the clear are ordered on stream s (synchronises s)
    // (partitions of 2^part_bits slots)
    hipError_t setup(const unsigned long long *out_off, const uint32_t *G, int nq, uint32_t nwg,
                     bool two_measures, hipStream_t s, uint32_t part_bits = kPartBits);
    // after the COUNT pass: the prefix sums (device only)
    hipError_t prefix(hipStream_t s);
    // the record total of the last COUNT pass (synchronises)
    hipError_t total(unsigned long long *out, hipStream_t s);
    hipError_t reserve(unsigned long long records);
    // first partition of query q
    uint32_t part_base(int q) const { return qbase_.empty() ? 0 : qbase_[(size_t)q]; }
    // (gcnt may be null: sums only)
    hipError_t reduce(unsigned long long *gsum, unsigned long long *gcnt, int num_cus,
                      hipStream_t s);
    StageArgs args(uint32_t cur_lds, uint32_t start_lds, unsigned long long *gsum, unsigned long long *gcnt) const;
    // LDS bytes of the counters/cursors plus the u64 range starts
    uint32_t lds_bytes() const;
    // records dropped because the buffers were too small (synchronises)
    hipError_t overflowed(uint32_t *out, hipStream_t s);

    // one-pass form: size the pool for `records` expected matches plus the
    // chunks `nwg` workgroups can leave open, with chunks of 2^ch_bits
    // records in format fmt (COMPACT: measures from mbase)
    hipError_t chunk_reserve(unsigned long long records, uint32_t nwg, uint32_t ch_bits, uint32_t fmt,
                             int32_t mbase);
    // clear the pool cursor and counters before a scan (kernels only)
    hipError_t chunk_clear(hipStream_t s);
    ChunkArgs chunk_args() const;
    // group the chunks by partition and reduce them into the outputs
    hipError_t chunk_reduce(unsigned long long *gsum, unsigned long long *gcnt, int num_cus,
                            hipStream_t s);
    void release();
    Stager() = default;
    Stager(const Stager &) = delete;
    Stager &operator=(const Stager &) = delete;
    ~Stager() { release(); }

    // one-pass scratch
    void *pool = nullptr;
    unsigned long long pool_cap = 0;      // records
    uint32_t pool_ch_bits = 0;
    uint32_t pool_fmt = REC_NARROW;
    int32_t pool_mbase = 0;
    unsigned long long *pool_cursor = nullptr;
    uint32_t *chunk_part = nullptr;
    uint32_t *chunk_fill = nullptr;
    unsigned long long *chunk_order = nullptr;  // (fill << 32) | chunk id, grouped by partition
    uint32_t *part_chunks = nullptr;      // [nparts + 1] chunk counts, then starts
    uint32_t *part_ccur = nullptr;        // [nparts] scatter cursors
    uint32_t *citem_begin = nullptr;      // [nparts + 1] reduce items per partition

  private:
    std::vector<uint32_t> qbase_;
    size_t cap_[8] = {};   // element capacities of the eight scratch arrays
    size_t ccap_[3] = {};  // capacities of part_chunks, part_ccur, citem_begin
};

}  // namespace olap
