// olap_kernels.hip -- gfx950 kernels of the star-join group-by path.
//
// Layout: fact columns are int32 arrays in HBM read with 128-bit loads (4
// rows per lane); dimension vectors are re-packed to 1/2/4/8/16/32-bit codes
// (all-ones = NULL) so the probed dimensions sit in LDS; group tables
// (u64 sum + u32 count per slot) sit in LDS when they fit and are flushed
// with one global atomic pair per non-empty slot, otherwise rows update the
// int64 output in HBM with global atomics. Compaction uses wave64 ballot +
// mbcnt with one LDS atomic per wave.
#include "olap_internal.h"

#include <type_traits>

namespace olap {

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ void atomic_add_u64_dev(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_u32_dev(uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_u64_lds(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t atomic_add_u32_lds(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

static inline int grid_for(int64_t work_items, int per_block, int cap)
{
    int64_t g = (work_items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ------------------------------------------------------------------------
// small utility kernels
// ------------------------------------------------------------------------

__global__ void k_zero_u64(unsigned long long *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_zero_u32(uint32_t *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// The outputs of a run and its error counter in one launch (a run's set-up
// is otherwise three dependent launches of a few microseconds each); sum may
// be null.
__global__ void k_zero_run(unsigned long long *sum, unsigned long long *count, size_t n, uint32_t *err)
{
    const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i0 == 0 && err) *err = 0u;
    for (size_t i = i0; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (sum) sum[i] = 0ull;
        count[i] = 0ull;
    }
}

hipError_t launch_zero_run(unsigned long long *sum, unsigned long long *count, size_t n, uint32_t *err,
                           hipStream_t s)
{
    k_zero_run<<<grid_for((int64_t)std::max<size_t>(n, 1), 256 * 4, 4096), 256, 0, s>>>(sum, count, n, err);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_zero_u64<<<grid_for((int64_t)n, 256 * 4, 4096), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_zero_u32<<<grid_for((int64_t)n, 256 * 4, 4096), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// synthetic data: SplitMix64 finaliser over (stream key + (row+1)*golden)
// ------------------------------------------------------------------------

__global__ void k_gen_uniform(int32_t *out, int64_t n, int64_t first, uint64_t key, int32_t lo,
                              uint64_t range)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t x = mix64(key + (uint64_t)(first + i + 1) * 0x9e3779b97f4a7c15ULL);
        out[i] = (int32_t)((int64_t)lo + (int64_t)(((x >> 32) * range) >> 32));
    }
}

__global__ void k_gen_zipf(int32_t *out, int64_t n, int64_t first, uint64_t key,
                           const uint64_t *thr, int64_t d)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t u = mix64(key + (uint64_t)(first + i + 1) * 0x9e3779b97f4a7c15ULL) >> 32;
        int64_t a = 0, b = d - 1;
        while (a < b) {
            int64_t m = (a + b) >> 1;
            if (thr[m] > u) b = m; else a = m + 1;
        }
        out[i] = (int32_t)a;
    }
}

hipError_t launch_gen_uniform(int32_t *out, int64_t n, int64_t first, uint64_t key, int32_t lo,
                              int32_t hi, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo);
    k_gen_uniform<<<grid_for(n, 256 * 8, 8192), 256, 0, s>>>(out, n, first, key, lo, range);
    return hipGetLastError();
}

hipError_t launch_gen_zipf(int32_t *out, int64_t n, int64_t first, uint64_t key,
                           const uint64_t *thr, int64_t d, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_gen_zipf<<<grid_for(n, 256 * 8, 8192), 256, 0, s>>>(out, n, first, key, thr, d);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// step 1: filter + group column -> dimension vector
// ------------------------------------------------------------------------

__global__ void k_build_dimvec(const int32_t *filter, const int32_t *group, int64_t rows,
                               int32_t lo, int32_t hi, int32_t bitmap, int32_t group_mod,
                               int32_t code_bytes, void *out)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        int32_t f = filter[r];
        bool pass = f >= lo && f < hi;
        int32_t code = 0;
        if (bitmap) {
            int32_t g = group[r];
            code = group_mod > 0 ? (int32_t)((uint32_t)g % (uint32_t)group_mod) : g;
        }
        if (code_bytes == 1)
            ((int8_t *)out)[r] = pass ? (int8_t)code : (int8_t)OLAP_DIM_NULL_I8;
        else
            ((int32_t *)out)[r] = pass ? code : OLAP_DIM_NULL_I32;
    }
}

hipError_t launch_build_dimvec(const int32_t *filter, const int32_t *group, int64_t rows,
                               int32_t lo, int32_t hi, int32_t bitmap, int32_t group_mod,
                               int32_t code_bytes, void *out, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    k_build_dimvec<<<grid_for(rows, 256, 4096), 256, 0, s>>>(filter, group, rows, lo, hi, bitmap,
                                                             group_mod, code_bytes, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// dimension-vector validation and packing
// ------------------------------------------------------------------------

__device__ __forceinline__ int32_t read_code(const void *dv, int32_t cb, int64_t r, bool &is_null)
{
    int32_t v;
    if (cb == 1) {
        v = ((const int8_t *)dv)[r];
        is_null = v == OLAP_DIM_NULL_I8;
    } else {
        v = ((const int32_t *)dv)[r];
        is_null = v == OLAP_DIM_NULL_I32;
    }
    return v;
}

__global__ void k_dim_stats(const void *dv, int32_t cb, int64_t rows, int32_t *stats)
{
    int32_t mx = -1;
    int32_t neg = 0;
    int32_t nonnull = 0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        bool nul;
        int32_t v = read_code(dv, cb, r, nul);
        if (!nul) {
            nonnull++;
            if (v < 0) neg++;
            else mx = v > mx ? v : mx;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        int32_t omx = __shfl_xor(mx, o);
        int32_t oneg = __shfl_xor(neg, o);
        int32_t onn = __shfl_xor(nonnull, o);
        mx = omx > mx ? omx : mx;
        neg += oneg;
        nonnull += onn;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&stats[0], mx);
        if (neg) atomicAdd(&stats[1], neg);
        if (nonnull) atomicAdd(&stats[2], nonnull);
    }
}

hipError_t launch_dim_stats(const void *dimvec, int32_t code_bytes, int64_t rows,
                            int32_t *stats, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    k_dim_stats<<<grid_for(rows, 256 * 4, 1024), 256, 0, s>>>(dimvec, code_bytes, rows, stats);
    return hipGetLastError();
}

__global__ void k_pack(const void *dv, int32_t cb, int64_t rows, uint32_t lw, uint32_t *words,
                       uint32_t nwords)
{
    const uint32_t bits = 1u << lw;
    const uint32_t per = 32u >> lw;
    const uint32_t mask = bits == 32 ? 0xffffffffu : ((1u << bits) - 1u);
    for (uint32_t wi = blockIdx.x * blockDim.x + threadIdx.x; wi < nwords;
         wi += gridDim.x * blockDim.x) {
        uint32_t w = 0;
        for (uint32_t k = 0; k < per; k++) {
            int64_t r = (int64_t)wi * per + k;
            uint32_t c = mask;
            if (r < rows) {
                bool nul;
                int32_t v = read_code(dv, cb, r, nul);
                c = nul ? mask : ((uint32_t)v & mask);
            }
            w |= c << (k << lw);
        }
        words[wi] = w;
    }
}

hipError_t launch_pack(const void *dimvec, int32_t code_bytes, int64_t rows, uint32_t lw,
                       uint32_t *words, uint32_t nwords, hipStream_t s)
{
    if (nwords == 0) return hipSuccess;
    k_pack<<<grid_for(nwords, 256, 2048), 256, 0, s>>>(dimvec, code_bytes, rows, lw, words,
                                                       nwords);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// scan building blocks
// ------------------------------------------------------------------------

// 4 consecutive rows starting at r; present = bit k set when row r+k exists
template <bool VEC>
__device__ __forceinline__ v4i load4(const int32_t *__restrict__ p, uint64_t r, uint64_t rows,
                                     uint32_t &present)
{
    v4i v = {0, 0, 0, 0};
    if (r + 4 <= rows) {
        present = 0xFu;
        if (VEC) {
            v = __builtin_nontemporal_load((const v4i *)(p + r));
        } else {
            v.x = p[r];
            v.y = p[r + 1];
            v.z = p[r + 2];
            v.w = p[r + 3];
        }
    } else {
        present = 0u;
        if (r < rows) { v.x = p[r]; present |= 1u; }
        if (r + 1 < rows) { v.y = p[r + 1]; present |= 2u; }
        if (r + 2 < rows) { v.z = p[r + 2]; present |= 4u; }
    }
    return v;
}

template <bool VEC>
__device__ __forceinline__ void store4(int32_t *__restrict__ p, uint64_t r, uint64_t rows, v4i v)
{
    if (r + 4 <= rows) {
        if (VEC) {
            *(v4i *)(p + r) = v;
        } else {
            p[r] = v.x;
            p[r + 1] = v.y;
            p[r + 2] = v.z;
            p[r + 3] = v.w;
        }
    } else {
        if (r < rows) p[r] = v.x;
        if (r + 1 < rows) p[r + 1] = v.y;
        if (r + 2 < rows) p[r + 2] = v.z;
    }
}

// probe one key: returns false when the key finds no row or a NULL code
// IN_LDS selects typed pointers: an LDS-resident vector is read with
// ds_read, which waits for LDS only (a generic pointer would make it a flat
// load, and a flat load's wait also waits for the next tile's prefetched
// columns)
template <bool IN_LDS>
__device__ __forceinline__ bool probe1(const StepDesc &st, const uint32_t *lds, int32_t key,
                                       uint32_t &code)
{
    const uint32_t idx = (uint32_t)key;
    const bool inb = idx < st.dim_rows;
    const uint32_t wi = inb ? (idx >> (5u - st.lw)) : 0u;
    const uint32_t word = IN_LDS ? ((const __attribute__((address_space(3))) uint32_t *)lds)[st.lds_off + wi]
                                 : ((const __attribute__((address_space(1))) uint32_t *)st.words)[wi];
    code = (word >> ((idx & ((32u >> st.lw) - 1u)) << st.lw)) & st.mask;
    return inb && code != st.mask;
}

__device__ __forceinline__ bool probe1_any(const StepDesc &st, const uint32_t *lds, int32_t key, uint32_t &code)
{
    return st.in_lds ? probe1<true>(st, lds, key, code) : probe1<false>(st, lds, key, code);
}

template <bool IN_LDS>
__device__ __forceinline__ void probe4_in(const StepDesc &st, const uint32_t *lds, v4i f,
                                          uint32_t (&gid)[4], uint32_t &alive, bool first)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if ((alive >> k) & 1u) {
            uint32_t code;
            bool ok = probe1<IN_LDS>(st, lds, f[k], code);
            const uint32_t add = code * st.factor;
            gid[k] = first ? add : gid[k] + add;
            if (!ok) alive &= ~(1u << k);
        }
    }
}

// probe 4 rows of a chunk; rows whose bit is clear in alive are skipped. The
// residency branch is wave-uniform and encloses the use of the loaded words,
// so the two paths meet with their waits done.
__device__ __forceinline__ void probe4(const StepDesc &st, const uint32_t *lds, v4i f,
                                       uint32_t (&gid)[4], uint32_t &alive, bool first)
{
    if (st.in_lds)
        probe4_in<true>(st, lds, f, gid, alive, first);
    else
        probe4_in<false>(st, lds, f, gid, alive, first);
}

// copy LDS-resident packed vectors; zero the LDS group table
template <int NDIM, bool GLDS>
__device__ __forceinline__ void lds_init(const ScanArgs &a, uint32_t *lds)
{
#pragma unroll
    for (int s = 0; s < NDIM; s++) {
        if (a.st[s].in_lds) {
            const uint32_t n = a.st[s].nwords;
            const uint32_t off = a.st[s].lds_off;
            const uint32_t *src = a.st[s].words;
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[off + i] = src[i];
        }
    }
    if (GLDS) {
        unsigned long long *lsum = (unsigned long long *)(lds + a.gt_off);
        uint32_t *lcnt = lds + a.gt_off + 2u * a.G;
        for (uint32_t g = threadIdx.x; g < a.G; g += blockDim.x) {
            lsum[g] = 0ull;
            lcnt[g] = 0u;
        }
    } else if (a.hot_bits) {
        const uint32_t S = 1u << a.hot_bits;
        uint32_t *tags = lds + a.hot_off;
        unsigned long long *hsum = (unsigned long long *)(tags + 2u * S);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
            tags[i] = kHotEmpty;
            tags[S + i] = 0u;
            hsum[i] = 0ull;
        }
    }
}

template <bool GLDS>
__device__ __forceinline__ void aggregate(const ScanArgs &a, uint32_t *lds, uint32_t gid,
                                          unsigned long long m)
{
    if (gid >= a.G) {
        atomic_add_u32_dev(a.err, 1u);
        return;
    }
    if (GLDS) {
        unsigned long long *lsum = (unsigned long long *)(lds + a.gt_off);
        uint32_t *lcnt = lds + a.gt_off + 2u * a.G;
        if (!a.count_only) atomic_add_u64_lds(&lsum[gid], m);
        atomic_add_u32_lds(&lcnt[gid], 1u);
    } else {
        if (a.hot_bits) {
            // the group's cache slot (multiplicative hash); an empty slot is
            // claimed by the first group that reaches it and keeps it
            const uint32_t S = 1u << a.hot_bits;
            uint32_t *tags = lds + a.hot_off;
            const uint32_t slot = (gid * 0x9e3779b1u) >> (32u - a.hot_bits);
            uint32_t tag = __hip_atomic_load(&tags[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (tag == kHotEmpty) {
                __hip_atomic_compare_exchange_strong(&tags[slot], &tag, gid, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                if (tag == kHotEmpty) tag = gid;  // claimed here
            }
            if (tag == gid) {
                if (!a.count_only) atomic_add_u64_lds((unsigned long long *)(tags + 2u * S) + slot, m);
                atomic_add_u32_lds(tags + S + slot, 1u);
                return;
            }
        }
        if (!a.count_only) atomic_add_u64_dev(&a.gsum[gid], m);
        atomic_add_u64_dev(&a.gcnt[gid], 1ull);
    }
}

template <bool GLDS>
__device__ __forceinline__ void flush(const ScanArgs &a, uint32_t *lds)
{
    if (!GLDS) {
        if (!a.hot_bits) return;
        // the cached groups go out with one atomic pair per slot
        __syncthreads();
        const uint32_t S = 1u << a.hot_bits;
        const uint32_t *tags = lds + a.hot_off;
        const unsigned long long *hsum = (const unsigned long long *)(tags + 2u * S);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
            const uint32_t c = tags[S + i];
            if (c) {
                atomic_add_u64_dev(&a.gcnt[tags[i]], (unsigned long long)c);
                if (!a.count_only) atomic_add_u64_dev(&a.gsum[tags[i]], hsum[i]);
            }
        }
        return;
    }
    __syncthreads();
    const unsigned long long *lsum = (const unsigned long long *)(lds + a.gt_off);
    const uint32_t *lcnt = lds + a.gt_off + 2u * a.G;
    for (uint32_t g = threadIdx.x; g < a.G; g += blockDim.x) {
        const uint32_t c = lcnt[g];
        if (c) {
            atomic_add_u64_dev(&a.gcnt[g], (unsigned long long)c);
            if (!a.count_only) atomic_add_u64_dev(&a.gsum[g], lsum[g]);
        }
    }
}

__device__ __forceinline__ unsigned long long measure4(const ScanArgs &a, v4i m1, v4i m2, int k)
{
    long long v = (long long)m1[k];
    if (a.m2) v += (long long)m2[k];
    return (unsigned long long)v;
}

// wave64 stream compaction: returns this lane's slot when want is set
__device__ __forceinline__ uint32_t wave_alloc(uint32_t *counter, bool want)
{
    const unsigned long long mask = __ballot(want);
    const uint32_t before = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    uint32_t base = 0;
    if (mask) {
        const int leader = __ffsll(mask) - 1;
        if ((int)__lane_id() == leader) base = atomic_add_u32_lds(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
    }
    return base + before;
}

// ------------------------------------------------------------------------
// row-wise fused scan (OLAPcore_row_thread semantics): probe steps in plan
// order, per-row early exit; a later step's 128-bit load is issued only for
// chunks that still hold a live row, so dead lines are never fetched.
// ------------------------------------------------------------------------

// one match of a staged scan: count it, or write its record
template <int MODE>
__device__ __forceinline__ void emit_staged(const ScanArgs &a, uint32_t *lds, uint32_t gid,
                                            int32_t m1, int32_t m2)
{
    if (gid >= a.G) {
        atomic_add_u32_dev(a.err, 1u);
        return;
    }
    if (MODE == AGG_COUNT) stage_count(a.stage, lds, gid >> kPartBits);
    else stage_scatter(a.stage, lds, gid >> kPartBits, gid & (kPartSlots - 1u), m1, m2);
}

// one match of a one-pass staged scan: its record goes to its partition's
// current chunk, or to the spare once the current one is full (a tile adds
// at most one chunk of records to a partition whose fill is below a chunk;
// a full chunk retires after the tile); with no chunk (the pool is used up)
// it is added to the outputs directly
__device__ __forceinline__ void emit_chunk(const ScanArgs &a, uint32_t *lds, uint32_t gid, int32_t m1, int32_t m2)
{
    if (gid >= a.G) {
        atomic_add_u32_dev(a.err, 1u);
        return;
    }
    const ChunkArgs &c = a.chunk;
    const uint32_t np = c.nparts;
    const uint32_t CH = 1u << c.ch_bits;
    const uint32_t part = gid >> kPartBits;
    const uint32_t glow = gid & (kPartSlots - 1u);
    stage_l_u32 *pc_cur = (stage_l_u32 *)lds + a.pc_lds;
    stage_l_u32 *pc_fill = pc_cur + 2u * np;
    const uint32_t pos = __hip_atomic_fetch_add(pc_fill + part, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t id = pos < CH ? pc_cur[part] : pc_cur[np + part];
    const uint32_t off = pos < CH ? pos : pos - CH;
    if (id == kNoChunk) {
        const unsigned long long o = c.part_out[part] + glow;
        atomic_add_u64_dev(&a.gsum[o], (unsigned long long)((long long)m1 + (long long)m2));
        atomic_add_u64_dev(&a.gcnt[o], 1ull);
        atomic_add_u32_dev(c.overflow, 1u);
        return;
    }
    const unsigned long long base = (unsigned long long)id << c.ch_bits;
    if (c.fmt == REC_WIDE) {
        ((uint4 *)c.pool)[base + off] = make_uint4(glow, (uint32_t)m1, (uint32_t)m2, 0u);
    } else {
        // a chunk holds CH u32 measures, then CH u16 slots
        char *cp = (char *)c.pool + 6ull * base;
        ((uint32_t *)cp)[off] = (uint32_t)m1;
        ((uint16_t *)(cp + 4u * CH))[off] = (uint16_t)glow;
    }
}

// one-pass staging, after a tile (all threads, between two barriers): full
// chunks retire, the spare becomes the current chunk and a new spare is
// claimed
__device__ __forceinline__ void retire_chunks(const ScanArgs &a, uint32_t *lds)
{
    const ChunkArgs &c = a.chunk;
    const uint32_t np = c.nparts;
    const uint32_t CH = 1u << c.ch_bits;
    stage_l_u32 *pc_cur = (stage_l_u32 *)lds + a.pc_lds;
    stage_l_u32 *pc_spare = pc_cur + np;
    stage_l_u32 *pc_fill = pc_spare + np;
    for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
        const uint32_t f = pc_fill[p];
        if (f >= CH) {
            const uint32_t cur = pc_cur[p];
            if (cur != kNoChunk) c.chunk_fill[cur] = CH;
            pc_cur[p] = pc_spare[p];
            pc_fill[p] = f - CH;
            pc_spare[p] = chunk_claim(c, p);
        }
    }
}

// MODE: AGG_DIRECT aggregates into the LDS table (GLDS) or with global
// atomics; AGG_COUNT / AGG_SCATTER are the two scan passes of the staged
// aggregation (the COUNT pass never reads the measures); AGG_CHUNK is the
// one-pass staged aggregation (records into pool chunks)
template <int NDIM, bool VEC, bool GLDS, int BS, int U, int MODE = AGG_DIRECT, bool EAGER = false>
__global__ __launch_bounds__(BS) void k_rowwise(ScanArgs a)
{
    constexpr uint64_t TILE = (uint64_t)BS * 4u * U;
    extern __shared__ __align__(16) uint32_t lds[];
    lds_init<NDIM, GLDS>(a, lds);
    stage_begin<MODE>(a.stage, lds);
    if (MODE == AGG_CHUNK) {
        stage_l_u32 *pc = (stage_l_u32 *)lds + a.pc_lds;
        const uint32_t np = a.chunk.nparts;
        for (uint32_t p = threadIdx.x; p < np; p += BS) {
            pc[p] = chunk_claim(a.chunk, p);
            pc[np + p] = chunk_claim(a.chunk, p);
            pc[2u * np + p] = 0u;
        }
    }
    __syncthreads();
    const uint64_t rows = a.rows;
    const uint32_t tid = threadIdx.x;
    // software pipeline: the first key column of the next tile is in flight
    // while the dependent probe chain of the current tile runs
    v4i nf[U], nf1[U];
    uint32_t np[U];
    uint64_t tile = blockIdx.x;
    constexpr bool eager = NDIM > 1 && EAGER;
    if (tile < a.ntiles) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t r = tile * TILE + (uint64_t)(u * BS + tid) * 4u;
            nf[u] = load4<VEC>(a.st[0].fk, r, rows, np[u]);
            if (eager) {
                uint32_t dummy;
                nf1[u] = load4<VEC>(a.st[NDIM > 1 ? 1 : 0].fk, r, rows, dummy);
            }
        }
    }
    for (; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile * TILE;
        v4i f[U], f1[U];
        uint32_t alive[U];
        uint32_t gid[U][4];
#pragma unroll
        for (int u = 0; u < U; u++) {
            f[u] = nf[u];
            f1[u] = nf1[u];
            alive[u] = np[u];
        }
        const uint64_t next = tile + gridDim.x;
        if (next < a.ntiles) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t r = next * TILE + (uint64_t)(u * BS + tid) * 4u;
                nf[u] = load4<VEC>(a.st[0].fk, r, rows, np[u]);
                if (eager) {
                    uint32_t dummy;
                    nf1[u] = load4<VEC>(a.st[NDIM > 1 ? 1 : 0].fk, r, rows, dummy);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) probe4(a.st[0], lds, f[u], gid[u], alive[u], true);
#pragma unroll
        for (int s = 1; s < NDIM; s++) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (s == 1 && eager) {
                    f[u] = f1[u];
                } else if (alive[u]) {
                    uint32_t dummy;
                    const uint64_t r = base + (uint64_t)(u * BS + tid) * 4u;
                    f[u] = load4<VEC>(a.st[s].fk, r, rows, dummy);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (alive[u]) probe4(a.st[s], lds, f[u], gid[u], alive[u], false);
        }
        v4i m1[U], m2[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            m1[u] = v4i{0, 0, 0, 0};
            m2[u] = v4i{0, 0, 0, 0};
            if (MODE != AGG_COUNT && alive[u] && !a.count_only) {
                uint32_t dummy;
                const uint64_t r = base + (uint64_t)(u * BS + tid) * 4u;
                m1[u] = load4<VEC>(a.m1, r, rows, dummy);
                if (a.m2) m2[u] = load4<VEC>(a.m2, r, rows, dummy);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (alive[u]) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if ((alive[u] >> k) & 1u) {
                        if (MODE == AGG_DIRECT)
                            aggregate<GLDS>(a, lds, gid[u][k], measure4(a, m1[u], m2[u], k));
                        else if (MODE == AGG_CHUNK)
                            emit_chunk(a, lds, gid[u][k], m1[u][k], a.m2 ? m2[u][k] : 0);
                        else
                            emit_staged<MODE>(a, lds, gid[u][k], m1[u][k], a.m2 ? m2[u][k] : 0);
                    }
            }
        }
        if (MODE == AGG_CHUNK) {
            __syncthreads();
            retire_chunks(a, lds);
            __syncthreads();
        }
    }
    if (MODE == AGG_CHUNK) {
        // fills of the chunks still open (below a chunk after the last retire)
        const uint32_t np = a.chunk.nparts;
        const stage_l_u32 *pc = (const stage_l_u32 *)lds + a.pc_lds;
        for (uint32_t p = tid; p < np; p += BS) {
            if (pc[p] != kNoChunk) a.chunk.chunk_fill[pc[p]] = pc[2u * np + p];
            if (pc[np + p] != kNoChunk) a.chunk.chunk_fill[pc[np + p]] = 0u;
        }
    }
    if (MODE == AGG_DIRECT) {
        flush<GLDS>(a, lds);
    } else if (MODE == AGG_COUNT) {
        __syncthreads();
        stage_end<MODE>(a.stage, lds);
    }
}

// ------------------------------------------------------------------------
// row-wise scan with every dependent fetch one iteration ahead (direct
// aggregation). Iteration i probes the streamed first step(s) of the
// workgroup's i-th tile and issues the next step's key loads for its live
// chunks; that tile's later steps and its measures follow in iterations
// i+1, i+2, ..., each probing what the previous iteration loaded and issuing
// the next loads. So every load is in flight for a whole iteration, and a
// wave waits once per iteration instead of once per step. The loads are
// buffer loads over the tile: a dead chunk's lane carries an offset past the
// buffer's end and reads 0 without touching memory, so every load
// instruction issues unconditionally and the compiler's waits stay exact.
// Full tiles only; the rows after the last full tile take a plain loop.
// ------------------------------------------------------------------------

constexpr int kBufWord3 = 0x00020000;      // buffer resource word 3 on gfx9: 32-bit data
constexpr uint32_t kDeadOff = 0x40000000u;  // past the end of any tile's buffer
constexpr int kBufNT = 2;                   // nontemporal: the columns stream once

__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_buf(const int32_t *col, uint64_t row, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)(col + row), (short)0, (int)bytes, kBufWord3);
}

__device__ __forceinline__ v4i buf_load4(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return (v4i)__builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, kBufNT);
}

template <int NDIM, bool GLDS, int BS, int U, bool EAGER>
__global__ __launch_bounds__(BS) void k_rowwise_pipe(ScanArgs a)
{
    constexpr uint64_t TILE = (uint64_t)BS * 4u * U;
    constexpr uint32_t TB = (uint32_t)TILE * 4u;       // bytes of one column of a tile
    constexpr int FIRST = NDIM > 1 && EAGER ? 2 : 1;    // steps probed from the streamed keys
    constexpr int D = NDIM - FIRST + 1;                 // stages: the later steps, then the measures
    extern __shared__ __align__(16) uint32_t lds[];
    lds_init<NDIM, GLDS>(a, lds);
    __syncthreads();
    const uint32_t tid = threadIdx.x;
    const uint64_t full = a.rows / TILE;
    const uint64_t mine = blockIdx.x < full ? (full - blockIdx.x + gridDim.x - 1) / gridDim.x : 0;
    const bool load_m = !a.count_only;
    const bool has_m2 = a.m2 != nullptr;
    uint32_t off[U];
#pragma unroll
    for (int u = 0; u < U; u++) off[u] = (uint32_t)(u * BS + tid) * 16u;

    // the streamed keys of the next tile
    v4i nf0[U], nf1[U];
    auto issue_first = [&](uint64_t i) {
        const bool ok = i < mine;
        const uint64_t row = ok ? (blockIdx.x + i * gridDim.x) * TILE : 0;
        const __amdgpu_buffer_rsrc_t r0 = tile_buf(a.st[0].fk, row, ok ? TB : 0u);
#pragma unroll
        for (int u = 0; u < U; u++) nf0[u] = buf_load4(r0, off[u]);
        if (FIRST == 2) {
            const __amdgpu_buffer_rsrc_t r1 = tile_buf(a.st[NDIM > 1 ? 1 : 0].fk, row, ok ? TB : 0u);
#pragma unroll
            for (int u = 0; u < U; u++) nf1[u] = buf_load4(r1, off[u]);
        }
    };

    // stage s holds a tile whose steps before FIRST + s are probed: its live
    // rows, partial gids, first row, and the loads in flight for it (the keys
    // of step FIRST + s, or the measures in the last stage)
    v4i ld[D][U], ld2[U];
    uint32_t al[D][U];
    uint32_t gd[D][U][4];
    uint64_t sb[D];
#pragma unroll
    for (int s = 0; s < D; s++) {
        sb[s] = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            ld[s][u] = v4i{0, 0, 0, 0};
            al[s][u] = 0u;
            gd[s][u][0] = gd[s][u][1] = gd[s][u][2] = gd[s][u][3] = 0u;
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) ld2[u] = v4i{0, 0, 0, 0};

    issue_first(0);
    for (uint64_t it = 0; it < mine + D; it++) {
        // 1. the oldest stage's measures have arrived: aggregate its matches
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (al[D - 1][u]) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if ((al[D - 1][u] >> k) & 1u) {
                        long long m = (long long)ld[D - 1][u][k];
                        if (has_m2) m += (long long)ld2[u][k];
                        aggregate<GLDS>(a, lds, gd[D - 1][u][k], (unsigned long long)m);
                    }
            }
        }
        // 2. probe what the other stages loaded, each moving one stage on
#pragma unroll
        for (int s = D - 2; s >= 0; s--) {
            sb[s + 1] = sb[s];
#pragma unroll
            for (int u = 0; u < U; u++) {
                uint32_t alive = al[s][u];
                uint32_t g[4] = {gd[s][u][0], gd[s][u][1], gd[s][u][2], gd[s][u][3]};
                if (alive) probe4(a.st[FIRST + s], lds, ld[s][u], g, alive, false);
                al[s + 1][u] = alive;
#pragma unroll
                for (int k = 0; k < 4; k++) gd[s + 1][u][k] = g[k];
            }
        }
        // ... and the new tile's streamed steps
        const bool ok = it < mine;
        sb[0] = ok ? (blockIdx.x + it * gridDim.x) * TILE : 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint32_t alive = ok ? 0xFu : 0u;
            uint32_t g[4] = {0u, 0u, 0u, 0u};
            probe4(a.st[0], lds, nf0[u], g, alive, true);
            if (FIRST == 2 && alive) probe4(a.st[NDIM > 1 ? 1 : 0], lds, nf1[u], g, alive, false);
            al[0][u] = alive;
#pragma unroll
            for (int k = 0; k < 4; k++) gd[0][u][k] = g[k];
        }
        // 3. issue this iteration's loads in the order the next one uses
        // them, then the next tile's streamed keys
#pragma unroll
        for (int s = D - 1; s >= 0; s--) {
            if (s == D - 1) {
                const __amdgpu_buffer_rsrc_t r1 = tile_buf(a.m1, sb[s], TB);
#pragma unroll
                for (int u = 0; u < U; u++) ld[s][u] = buf_load4(r1, al[s][u] && load_m ? off[u] : kDeadOff);
                if (has_m2) {
                    const __amdgpu_buffer_rsrc_t r2 = tile_buf(a.m2, sb[s], TB);
#pragma unroll
                    for (int u = 0; u < U; u++) ld2[u] = buf_load4(r2, al[s][u] && load_m ? off[u] : kDeadOff);
                }
            } else {
                const __amdgpu_buffer_rsrc_t rk = tile_buf(a.st[FIRST + s].fk, sb[s], TB);
#pragma unroll
                for (int u = 0; u < U; u++) ld[s][u] = buf_load4(rk, al[s][u] ? off[u] : kDeadOff);
            }
        }
        issue_first(it + 1);
    }

    // rows after the last full tile: every step in turn
    for (uint64_t r = full * TILE + ((uint64_t)blockIdx.x * BS + tid) * 4u; r < a.rows;
         r += (uint64_t)gridDim.x * BS * 4u) {
        uint32_t alive, dummy;
        uint32_t g[4] = {0u, 0u, 0u, 0u};
        v4i f = load4<true>(a.st[0].fk, r, a.rows, alive);
        probe4(a.st[0], lds, f, g, alive, true);
#pragma unroll
        for (int s = 1; s < NDIM; s++) {
            if (!alive) break;
            f = load4<true>(a.st[s].fk, r, a.rows, dummy);
            probe4(a.st[s], lds, f, g, alive, false);
        }
        if (!alive) continue;
        v4i m1 = v4i{0, 0, 0, 0}, m2 = v4i{0, 0, 0, 0};
        if (load_m) {
            m1 = load4<true>(a.m1, r, a.rows, dummy);
            if (has_m2) m2 = load4<true>(a.m2, r, a.rows, dummy);
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
            if ((alive >> k) & 1u) aggregate<GLDS>(a, lds, g[k], measure4(a, m1, m2, k));
    }
    flush<GLDS>(a, lds);
}

// ------------------------------------------------------------------------
// row-wise scan with a survivor queue (direct aggregation). The streamed
// first step(s) are loaded PF tiles ahead and probed as they arrive; the rows
// that pass them go, with their partial gids, into the wave's LDS queue.
// Once the queue holds 64 rows, each lane takes one and runs the later steps
// and the measures for it with 4-byte gathers. When few rows pass the
// streamed steps, most tiles make no dependent fetch at all, and a wave
// waits only for keys it issued PF tiles earlier. Rows are numbered in 32
// bits (the host keeps larger tables off this kernel).
// ------------------------------------------------------------------------

// queue entries per wave when a drain takes B rows per lane: below 64 B
// before a chunk's push, which adds at most 256
constexpr uint32_t queue_cap(int B) { return 64u * (uint32_t)B + 256u; }

uint32_t queue_lds_bytes(int bs, int B) { return (uint32_t)(bs / 64) * 2u * queue_cap(B) * 4u; }

template <int NDIM, bool GLDS, int BS, int U, bool EAGER, int PF, int B>
__global__ __launch_bounds__(BS) void k_rowwise_queue(ScanArgs a)
{
    constexpr uint32_t QCAP = queue_cap(B);
    constexpr uint32_t BATCH = 64u * B;
    constexpr uint64_t TILE = (uint64_t)BS * 4u * U;
    constexpr uint32_t TB = (uint32_t)TILE * 4u;
    constexpr int FIRST = NDIM > 1 && EAGER ? 2 : 1;
    typedef __attribute__((address_space(3))) uint32_t lds_u32;
    extern __shared__ __align__(16) uint32_t lds[];
    lds_init<NDIM, GLDS>(a, lds);
    __syncthreads();
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    lds_u32 *qrow = (lds_u32 *)lds + a.q_off + (tid >> 6) * (2u * QCAP);
    lds_u32 *qgid = qrow + QCAP;
    const uint64_t full = a.rows / TILE;
    const uint64_t mine = blockIdx.x < full ? (full - blockIdx.x + gridDim.x - 1) / gridDim.x : 0;
    const bool load_m = !a.count_only;
    const bool has_m2 = a.m2 != nullptr;
    uint32_t off[U];
#pragma unroll
    for (int u = 0; u < U; u++) off[u] = (uint32_t)(u * BS + tid) * 16u;

    // the streamed keys of the next PF tiles: tile i goes to slot i % PF
    v4i nf0[PF][U], nf1[PF][U];
    auto issue_first = [&](auto slot, uint64_t i) {
        constexpr int P = decltype(slot)::value;
        const bool ok = i < mine;
        const uint64_t row = ok ? (blockIdx.x + i * gridDim.x) * TILE : 0;
        const __amdgpu_buffer_rsrc_t r0 = tile_buf(a.st[0].fk, row, ok ? TB : 0u);
#pragma unroll
        for (int u = 0; u < U; u++) nf0[P][u] = buf_load4(r0, off[u]);
        if (FIRST == 2) {
            const __amdgpu_buffer_rsrc_t r1 = tile_buf(a.st[NDIM > 1 ? 1 : 0].fk, row, ok ? TB : 0u);
#pragma unroll
            for (int u = 0; u < U; u++) nf1[P][u] = buf_load4(r1, off[u]);
        }
    };

    // the later steps and the measures of the n (<= 64 B) queued rows from
    // entry base: B rows per lane, their gathers of a step issued together
    auto drain = [&](uint32_t base, uint32_t n) {
        uint32_t row[B], g[B];
        bool alive[B];
#pragma unroll
        for (int j = 0; j < B; j++) {
            const uint32_t e = lane + 64u * (uint32_t)j;
            alive[j] = e < n;
            row[j] = alive[j] ? qrow[base + e] : 0u;
            g[j] = alive[j] ? qgid[base + e] : 0u;
        }
#pragma unroll
        for (int s = FIRST; s < NDIM; s++) {
            int32_t key[B];
#pragma unroll
            for (int j = 0; j < B; j++) key[j] = alive[j] ? a.st[s].fk[row[j]] : 0;
#pragma unroll
            for (int j = 0; j < B; j++) {
                if (alive[j]) {
                    uint32_t code;
                    alive[j] = probe1_any(a.st[s], lds, key[j], code);
                    g[j] += code * a.st[s].factor;
                }
            }
        }
        long long m[B];
#pragma unroll
        for (int j = 0; j < B; j++) {
            m[j] = 0;
            if (alive[j] && load_m) {
                m[j] = (long long)a.m1[row[j]];
                if (has_m2) m[j] += (long long)a.m2[row[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < B; j++)
            if (alive[j]) aggregate<GLDS>(a, lds, g[j], (unsigned long long)m[j]);
    };

    uint32_t qn = 0;  // queued rows of this wave (wave-uniform)
    auto iteration = [&](auto slot, uint64_t it) {
        constexpr int P = decltype(slot)::value;
        const bool ok = it < mine;
        const uint64_t tile_row = ok ? (blockIdx.x + it * gridDim.x) * TILE : 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint32_t alive = ok ? 0xFu : 0u;
            uint32_t g[4] = {0u, 0u, 0u, 0u};
            probe4(a.st[0], lds, nf0[P][u], g, alive, true);
            if (FIRST == 2 && alive) probe4(a.st[NDIM > 1 ? 1 : 0], lds, nf1[P][u], g, alive, false);
            const uint32_t row0 = (uint32_t)(tile_row + (uint64_t)(u * BS + tid) * 4u);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool want = (alive >> k) & 1u;
                const unsigned long long mask = __ballot(want);
                if (mask) {
                    const uint32_t before = __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                    if (want) {
                        qrow[qn + before] = row0 + (uint32_t)k;
                        qgid[qn + before] = g[k];
                    }
                    qn += (uint32_t)__popcll(mask);
                }
            }
            while (qn >= BATCH) {
                qn -= BATCH;
                drain(qn, BATCH);
            }
        }
        issue_first(slot, it + PF);
    };

    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, (PF > 1 ? 1 : 0)>;
    using S2 = std::integral_constant<int, (PF > 2 ? 2 : 0)>;
    using S3 = std::integral_constant<int, (PF > 3 ? 3 : 0)>;
    issue_first(S0{}, 0);
    if constexpr (PF > 1) issue_first(S1{}, 1);
    if constexpr (PF > 2) issue_first(S2{}, 2);
    if constexpr (PF > 3) issue_first(S3{}, 3);
    // the iterations past the workgroup's last tile have no live row and load
    // nothing, so the count rounds up to whole groups of PF
    const uint64_t iters = (mine + PF - 1) / PF * PF;
    for (uint64_t it = 0; it < iters; it += PF) {
        iteration(S0{}, it);
        if constexpr (PF > 1) iteration(S1{}, it + 1);
        if constexpr (PF > 2) iteration(S2{}, it + 2);
        if constexpr (PF > 3) iteration(S3{}, it + 3);
    }
    if (qn) drain(0u, qn);

    // rows after the last full tile: every step in turn
    for (uint64_t r = full * TILE + ((uint64_t)blockIdx.x * BS + tid) * 4u; r < a.rows;
         r += (uint64_t)gridDim.x * BS * 4u) {
        uint32_t alive, dummy;
        uint32_t g[4] = {0u, 0u, 0u, 0u};
        v4i f = load4<true>(a.st[0].fk, r, a.rows, alive);
        probe4(a.st[0], lds, f, g, alive, true);
#pragma unroll
        for (int s = 1; s < NDIM; s++) {
            if (!alive) break;
            f = load4<true>(a.st[s].fk, r, a.rows, dummy);
            probe4(a.st[s], lds, f, g, alive, false);
        }
        if (!alive) continue;
        v4i m1 = v4i{0, 0, 0, 0}, m2 = v4i{0, 0, 0, 0};
        if (load_m) {
            m1 = load4<true>(a.m1, r, a.rows, dummy);
            if (has_m2) m2 = load4<true>(a.m2, r, a.rows, dummy);
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
            if ((alive >> k) & 1u) aggregate<GLDS>(a, lds, g[k], measure4(a, m1, m2, k));
    }
    flush<GLDS>(a, lds);
}

// ------------------------------------------------------------------------
// vector-wise static vector (OLAPcore_vwm_sv_thread semantics): every step is
// probed for every row; all columns of a tile are loaded up front.
// ------------------------------------------------------------------------

template <int NDIM, bool VEC, bool GLDS>
__global__ __launch_bounds__(kBlock) void k_vwm_sv(ScanArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];
    lds_init<NDIM, GLDS>(a, lds);
    __syncthreads();
    const uint64_t rows = a.rows;
    const uint32_t tid = threadIdx.x;
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile * (uint64_t)kTileRW;
        v4i f[NDIM][kUnroll], m1[kUnroll], m2[kUnroll];
        uint32_t present[kUnroll], alive[kUnroll];
        uint32_t gid[kUnroll][4];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint64_t r = base + (uint64_t)(u * kBlock + tid) * 4u;
            uint32_t dummy;
#pragma unroll
            for (int s = 0; s < NDIM; s++) f[s][u] = load4<VEC>(a.st[s].fk, r, rows, present[u]);
            m1[u] = v4i{0, 0, 0, 0};
            m2[u] = v4i{0, 0, 0, 0};
            if (!a.count_only) {
                m1[u] = load4<VEC>(a.m1, r, rows, dummy);
                if (a.m2) m2[u] = load4<VEC>(a.m2, r, rows, dummy);
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            // static vector: every row is probed in every step (no early exit)
            uint32_t live = present[u];
#pragma unroll
            for (int k = 0; k < 4; k++) gid[u][k] = 0u;
#pragma unroll
            for (int s = 0; s < NDIM; s++) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if ((present[u] >> k) & 1u) {
                        uint32_t code;
                        bool ok = probe1_any(a.st[s], lds, f[s][u][k], code);
                        gid[u][k] += code * a.st[s].factor;
                        if (!ok) live &= ~(1u << k);
                    }
                }
            }
            alive[u] = live;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (alive[u]) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if ((alive[u] >> k) & 1u)
                        aggregate<GLDS>(a, lds, gid[u][k], measure4(a, m1[u], m2[u], k));
            }
        }
    }
    flush<GLDS>(a, lds);
}

// ------------------------------------------------------------------------
// vector-wise dynamic vector (OLAPcore_vwm_dv_thread semantics): per tile of
// kTileV rows, step 0 compacts (row offset, gid) of live rows into an LDS
// list with wave ballots; each later step gathers its key only for listed
// rows and re-compacts in place; the measure is gathered for the survivors.
// ------------------------------------------------------------------------

template <int NDIM, bool VEC, bool GLDS>
__global__ __launch_bounds__(kBlock) void k_vwm_dv(ScanArgs a)
{
    // Each wave owns a block of kTileV / kWaves rows of the tile and its own
    // slice of the LDS list: step 0 compacts the block's passing rows (row
    // offset, partial gid) with ballot / mbcnt, every later step re-filters
    // the list in place, and the survivors are aggregated. A wave's list is
    // its own, so the tile loop has no workgroup barrier.
    constexpr uint32_t kWaves = kBlock / 64u;
    constexpr uint32_t kWaveRows = (uint32_t)kTileV / kWaves;  // rows, and list entries, of a wave
    constexpr int kU = (int)(kWaveRows / 256u);                 // 128-bit chunks per lane
    extern __shared__ __align__(16) uint32_t lds[];
    lds_init<NDIM, GLDS>(a, lds);
    __syncthreads();
    const uint64_t rows = a.rows;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    typedef __attribute__((address_space(3))) uint32_t l32;
    typedef __attribute__((address_space(3))) uint16_t l16;
    l32 *lgid = (l32 *)(lds + a.list_off) + wave * kWaveRows;
    l16 *loff = (l16 *)((l32 *)(lds + a.list_off) + kTileV) + wave * kWaveRows;
    auto row_of = [&](int u, int k) { return wave * kWaveRows + (uint32_t)(u * 64 + (int)lane) * 4u + (uint32_t)k; };
    // software pipeline: the next tile's first key column is in flight while
    // the current tile's list is built
    v4i nf[kU];
    uint32_t nalive[kU];
    if (blockIdx.x < a.ntiles) {
#pragma unroll
        for (int u = 0; u < kU; u++)
            nf[u] = load4<VEC>(a.st[0].fk, (uint64_t)blockIdx.x * kTileV + row_of(u, 0), rows, nalive[u]);
    }
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile * (uint64_t)kTileV;
        v4i f[kU];
        uint32_t alive[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            f[u] = nf[u];
            alive[u] = nalive[u];
        }
        const uint64_t next = tile + gridDim.x;
        if (next < a.ntiles) {
#pragma unroll
            for (int u = 0; u < kU; u++) nf[u] = load4<VEC>(a.st[0].fk, next * kTileV + row_of(u, 0), rows, nalive[u]);
        }
        // step 0: compact the passing rows into the wave's list
        uint32_t n = 0u;
#pragma unroll
        for (int u = 0; u < kU; u++) {
            uint32_t gid[4];
            probe4(a.st[0], lds, f[u], gid, alive[u], true);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool want = (alive[u] >> k) & 1u;
                const unsigned long long m = __ballot(want);
                const uint32_t pos = n + (uint32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (want) {
                    loff[pos] = (uint16_t)row_of(u, k);
                    lgid[pos] = gid[k];
                }
                n += (uint32_t)__popcll(m);
            }
        }
        // later steps: re-filter the list in place, 64 entries at a time (a
        // round reads its entries before it writes, and writes only at or
        // below them)
#pragma unroll
        for (int s = 1; s < NDIM; s++) {
            uint32_t kept = 0u;
            for (uint32_t e0 = 0; e0 < n; e0 += 64u) {
                const uint32_t e = e0 + lane;
                const bool in = e < n;
                const uint32_t off = in ? (uint32_t)loff[e] : 0u;
                uint32_t g = in ? lgid[e] : 0u;
                const int32_t key = in ? a.st[s].fk[base + off] : 0;
                uint32_t code = 0u;
                const bool pass = in && probe1_any(a.st[s], lds, key, code);
                g += code * a.st[s].factor;
                const unsigned long long m = __ballot(pass);
                const uint32_t pos = kept + (uint32_t)__builtin_amdgcn_mbcnt_hi(
                                                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (pass) {
                    loff[pos] = (uint16_t)off;
                    lgid[pos] = g;
                }
                kept += (uint32_t)__popcll(m);
            }
            n = kept;
        }
        // the survivors' measures, then their aggregation
        for (uint32_t e0 = 0; e0 < n; e0 += 64u) {
            const uint32_t e = e0 + lane;
            if (e < n) {
                const uint64_t row = base + loff[e];
                unsigned long long m = 0ull;
                if (!a.count_only) {
                    long long v = (long long)a.m1[row];
                    if (a.m2) v += (long long)a.m2[row];
                    m = (unsigned long long)v;
                }
                aggregate<GLDS>(a, lds, lgid[e], m);
            }
        }
    }
    flush<GLDS>(a, lds);
}

// ------------------------------------------------------------------------
// column-wise static vector (join_cwm_sv / agg_cwm_sv semantics): one pass
// per step over a full per-row gid vector in HBM, then one aggregation pass.
// ------------------------------------------------------------------------

template <bool VEC>
__global__ __launch_bounds__(kBlock) void k_cwm_sv_join(ScanArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];
    const StepDesc &st = a.st[a.step];
    if (st.in_lds)
        for (uint32_t i = threadIdx.x; i < st.nwords; i += kBlock) lds[st.lds_off + i] = st.words[i];
    __syncthreads();
    const uint64_t rows = a.rows;
    const uint64_t nchunks = (rows + 3) / 4;
    const bool first = a.step == 0;
    const bool last = a.last_step != 0;
    for (uint64_t c = blockIdx.x * (uint64_t)kBlock + threadIdx.x; c < nchunks;
         c += (uint64_t)gridDim.x * kBlock) {
        const uint64_t r = c * 4u;
        uint32_t present, dummy;
        v4i g = first ? v4i{0, 0, 0, 0} : load4<VEC>(a.gvec, r, rows, present);
        uint32_t alive;
        if (first) {
            alive = r + 4 <= rows ? 0xFu : ((1u << (uint32_t)(rows - r)) - 1u);
        } else {
            alive = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) alive |= (g[k] != OLAP_GID_NULL ? 1u : 0u) << k;
            alive &= present;
        }
        const uint32_t was_alive = alive;
        if (alive) {
            v4i f = load4<VEC>(st.fk, r, rows, dummy);
            uint32_t gg[4] = {(uint32_t)g.x, (uint32_t)g.y, (uint32_t)g.z, (uint32_t)g.w};
            probe4(st, lds, f, gg, alive, first);
            g = v4i{(int)gg[0], (int)gg[1], (int)gg[2], (int)gg[3]};
        }
        const int32_t nul = last ? a.gvec_null : OLAP_GID_NULL;
        // a chunk whose rows an earlier step had all nulled keeps its words
        // (most chunks after a selective first step), unless the last step
        // writes a different null
        if (!first && !was_alive && nul == OLAP_GID_NULL) continue;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (!((alive >> k) & 1u)) g[k] = nul;
        store4<VEC>(a.gvec, r, rows, g);
    }
}

template <bool VEC, bool GLDS>
__global__ __launch_bounds__(kBlock) void k_cwm_sv_agg(ScanArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];
    lds_init<0, GLDS>(a, lds);
    __syncthreads();
    const uint64_t rows = a.rows;
    const uint64_t nchunks = (rows + 3) / 4;
    for (uint64_t c = blockIdx.x * (uint64_t)kBlock + threadIdx.x; c < nchunks;
         c += (uint64_t)gridDim.x * kBlock) {
        const uint64_t r = c * 4u;
        uint32_t present, dummy;
        const v4i g = load4<VEC>(a.gvec, r, rows, present);
        uint32_t alive = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) alive |= (g[k] != OLAP_GID_NULL ? 1u : 0u) << k;
        alive &= present;
        if (alive) {
            v4i m1 = {0, 0, 0, 0}, m2 = {0, 0, 0, 0};
            if (!a.count_only) {
                m1 = load4<VEC>(a.m1, r, rows, dummy);
                if (a.m2) m2 = load4<VEC>(a.m2, r, rows, dummy);
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if ((alive >> k) & 1u) aggregate<GLDS>(a, lds, (uint32_t)g[k], measure4(a, m1, m2, k));
        }
    }
    flush<GLDS>(a, lds);
}

// ------------------------------------------------------------------------
// column-wise dynamic vector (join_cwm_dv / agg_cwm_dv semantics): the fact
// table is cut into kTileV-row partitions, each owning a list segment in HBM
// (row id, gid) and an entry count -- the per-thread OID/groupID/index of
// the reference, one per partition. Step 0 fills the segments, each later
// step re-compacts its segment in place, the aggregation pass reads them.
// ------------------------------------------------------------------------

template <bool VEC>
__global__ __launch_bounds__(kBlock) void k_cwm_dv_join0(ScanArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];
    const StepDesc &st = a.st[0];
    if (st.in_lds)
        for (uint32_t i = threadIdx.x; i < st.nwords; i += kBlock) lds[st.lds_off + i] = st.words[i];
    // two counters used by alternate tiles: a tile's count is published, and
    // its counter cleared, after the next tile's barrier, so one barrier a
    // tile suffices
    uint32_t *counter = lds + a.list_off;
    if (threadIdx.x < 2u) counter[threadIdx.x] = 0u;
    __syncthreads();
    const uint64_t rows = a.rows;
    const uint32_t tid = threadIdx.x;
    constexpr int kU = kTileV / (kBlock * 4);
    uint32_t par = 0u;
    uint64_t prev = ~0ull;
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile * (uint64_t)kTileV;
        uint32_t *soid = a.list_oid + base;
        int32_t *sgid = a.list_gid + base;
        v4i f[kU];
        uint32_t alive[kU];
#pragma unroll
        for (int u = 0; u < kU; u++)
            f[u] = load4<VEC>(st.fk, base + (uint64_t)(u * kBlock + tid) * 4u, rows, alive[u]);
        __syncthreads();  // the previous tile's list is complete
        if (tid == 0 && prev != ~0ull) {
            a.list_cnt[prev] = counter[par ^ 1u];
            counter[par ^ 1u] = 0u;
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const uint64_t r = base + (uint64_t)(u * kBlock + tid) * 4u;
            uint32_t gid[4];
            probe4(st, lds, f[u], gid, alive[u], true);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool want = (alive[u] >> k) & 1u;
                const uint32_t pos = wave_alloc(counter + par, want);
                if (want) {
                    soid[pos] = (uint32_t)(r + k);
                    sgid[pos] = (int32_t)gid[k];
                }
            }
        }
        prev = tile;
        par ^= 1u;
    }
    __syncthreads();
    if (tid == 0 && prev != ~0ull) a.list_cnt[prev] = counter[par ^ 1u];
}

__global__ __launch_bounds__(kBlock) void k_cwm_dv_join(ScanArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];
    const StepDesc &st = a.st[a.step];
    if (st.in_lds)
        for (uint32_t i = threadIdx.x; i < st.nwords; i += kBlock) lds[st.lds_off + i] = st.words[i];
    // two counters used by alternate tiles (as in k_cwm_dv_join0): one
    // barrier a tile, between reading a segment and compacting it in place
    uint32_t *counter = lds + a.list_off;
    if (threadIdx.x < 2u) counter[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t tid = threadIdx.x;
    uint32_t par = 0u;
    uint64_t prev = ~0ull;
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile * (uint64_t)kTileV;
        uint32_t *soid = a.list_oid + base;
        int32_t *sgid = a.list_gid + base;
        const uint32_t n = a.list_cnt[tile];
        uint32_t eoid[kPerThreadV], egid[kPerThreadV];
        uint32_t ok = 0;
#pragma unroll
        for (int j = 0; j < kPerThreadV; j++) {
            const uint32_t e = tid + j * kBlock;
            if (e < n) {
                eoid[j] = soid[e];
                egid[j] = (uint32_t)sgid[e];
                uint32_t code;
                if (probe1_any(st, lds, st.fk[eoid[j]], code)) {
                    egid[j] += code * st.factor;
                    ok |= 1u << j;
                }
            }
        }
        __syncthreads();  // every entry of the segment has been read; the previous tile's list is complete
        if (tid == 0 && prev != ~0ull) {
            a.list_cnt[prev] = counter[par ^ 1u];
            counter[par ^ 1u] = 0u;
        }
#pragma unroll
        for (int j = 0; j < kPerThreadV; j++) {
            const bool want = (ok >> j) & 1u;
            const uint32_t pos = wave_alloc(counter + par, want);
            if (want) {
                soid[pos] = eoid[j];
                sgid[pos] = (int32_t)egid[j];
            }
        }
        prev = tile;
        par ^= 1u;
    }
    __syncthreads();
    if (tid == 0 && prev != ~0ull) a.list_cnt[prev] = counter[par ^ 1u];
}

template <bool GLDS>
__global__ __launch_bounds__(kBlock) void k_cwm_dv_agg(ScanArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];
    lds_init<0, GLDS>(a, lds);
    __syncthreads();
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile * (uint64_t)kTileV;
        const uint32_t n = a.list_cnt[tile];
        for (uint32_t e = threadIdx.x; e < n; e += kBlock) {
            const uint32_t row = a.list_oid[base + e];
            unsigned long long m = 0ull;
            if (!a.count_only) {
                long long v = (long long)a.m1[row];
                if (a.m2) v += (long long)a.m2[row];
                m = (unsigned long long)v;
            }
            aggregate<GLDS>(a, lds, (uint32_t)a.list_gid[base + e], m);
        }
    }
    flush<GLDS>(a, lds);
}

// ------------------------------------------------------------------------
// launch helpers
// ------------------------------------------------------------------------

template <typename K>
static hipError_t launch_persistent(K kernel, ScanArgs &a, int num_cus, uint32_t dyn_lds,
                                    uint64_t work_units, hipStream_t s, int bs = kBlock,
                                    int max_per_cu = 0)
{
    if (dyn_lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds);
        if (e != hipSuccess) return e;
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bs, dyn_lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    if (max_per_cu > 0 && per_cu > max_per_cu) per_cu = max_per_cu;
    uint64_t grid = (uint64_t)per_cu * (uint64_t)num_cus;
    if (grid > work_units) grid = work_units;
    if (grid < 1) grid = 1;
    kernel<<<(unsigned)grid, bs, dyn_lds, s>>>(a);
    return hipGetLastError();
}

#define OLAP_DISPATCH_NDIM(KERNEL, VEC, GLDS)                                      \
    switch (ndims) {                                                             \
    case 1: return launch_persistent(KERNEL<1, VEC, GLDS>, a, num_cus, dyn_lds, a.ntiles, s); \
    case 2: return launch_persistent(KERNEL<2, VEC, GLDS>, a, num_cus, dyn_lds, a.ntiles, s); \
    case 3: return launch_persistent(KERNEL<3, VEC, GLDS>, a, num_cus, dyn_lds, a.ntiles, s); \
    default: return launch_persistent(KERNEL<4, VEC, GLDS>, a, num_cus, dyn_lds, a.ntiles, s); \
    }

#define OLAP_DISPATCH_ALL(KERNEL)                                                  \
    if (vec) {                                                                   \
        if (a.gt_lds) { OLAP_DISPATCH_NDIM(KERNEL, true, true) }                  \
        else { OLAP_DISPATCH_NDIM(KERNEL, true, false) }                          \
    } else {                                                                     \
        if (a.gt_lds) { OLAP_DISPATCH_NDIM(KERNEL, false, true) }                 \
        else { OLAP_DISPATCH_NDIM(KERNEL, false, false) }                         \
    }

// Row-wise launch shapes: (threads per workgroup, 128-bit chunks per thread
// per tile). The autotuner times them on a live query.
// Shapes 6 and 7 run the pipelined kernel (every dependent fetch one tile
// ahead), 8-15 the survivor-queue kernel (1024 threads, streamed keys 2
// tiles ahead, drains of 1 or 4 rows per lane); 12-15 never stream the
// second step's keys with the first's, whatever share of their lines hold
// a survivor (the queue's gathers then fetch only those lines).
constexpr ScanShape kRowShapes[] = {{512, 4, 0},  {512, 2, 0},  {1024, 2, 0}, {1024, 4, 0}, {256, 8, 0},  {512, 8, 0},
                                {1024, 2, 1}, {512, 2, 1},  {1024, 2, 3}, {1024, 2, 4}, {1024, 4, 3}, {1024, 4, 4},
                                {1024, 2, 5}, {1024, 2, 6}, {1024, 4, 5}, {1024, 4, 6}};
const int kNumRowShapes = (int)(sizeof(kRowShapes) / sizeof(kRowShapes[0]));
static_assert(kRowShapes[kDefaultShapeStreamed].bs == 1024 && kRowShapes[kDefaultShapeStreamed].unroll == 4 &&
                  kRowShapes[kDefaultShapeStreamed].pipe == 4,
              "default streamed shape");
static_assert(kRowShapes[kDefaultShapeUnstreamed].bs == 1024 && kRowShapes[kDefaultShapeUnstreamed].unroll == 4 &&
                  kRowShapes[kDefaultShapeUnstreamed].pipe == 6,
              "default unstreamed shape");

template <int NDIM, bool VEC, bool GLDS>
static hipError_t rowwise_shape(ScanArgs &a, int num_cus, uint32_t dyn_lds, int shape,
                                hipStream_t s)
{
    const ScanShape sh = kRowShapes[shape];
    a.ntiles = (a.rows + (uint64_t)sh.bs * 4u * sh.unroll - 1) / ((uint64_t)sh.bs * 4u * sh.unroll);
    if (a.ntiles == 0) return hipSuccess;
    if (VEC && sh.pipe >= 3) {
        // the survivor queues go after the rest of the kernel's LDS; a table
        // with more rows than 32-bit row numbers, or queues that do not fit,
        // take the pipelined kernel of the same size
        const uint32_t qoff = (dyn_lds + 15u) & ~15u;
        const int B = sh.pipe == 3 || sh.pipe == 5 ? 1 : 4;
        const uint32_t dyn = qoff + queue_lds_bytes(1024, B);
        if (a.rows <= 0xFFFFFFFFull && dyn <= kLdsBytes) {
            a.q_off = qoff / 4u;
            const bool eager = NDIM > 1 && a.eager1 && sh.pipe <= 4;
            a.launched = OLAP_PIPE_QUEUE | (eager ? OLAP_PIPE_STREAM2 : 0);
#define OLAP_QUEUE(UX, BX)                                                                                        \
    return eager ? launch_persistent(k_rowwise_queue<NDIM, GLDS, 1024, UX, true, 2, BX>, a, num_cus, dyn, a.ntiles, s, 1024) \
                 : launch_persistent(k_rowwise_queue<NDIM, GLDS, 1024, UX, false, 2, BX>, a, num_cus, dyn, a.ntiles, s, 1024)
            if (sh.unroll == 2) {
                if (B == 1) OLAP_QUEUE(2, 1);
                OLAP_QUEUE(2, 4);
            }
            if (B == 1) OLAP_QUEUE(4, 1);
            OLAP_QUEUE(4, 4);
#undef OLAP_QUEUE
        }
    }
    if (VEC && sh.pipe) {
        const bool eager = NDIM > 1 && a.eager1 && (sh.pipe == 1 || sh.pipe == 3 || sh.pipe == 4);
        a.launched = OLAP_PIPE_AHEAD | (eager ? OLAP_PIPE_STREAM2 : 0);
#define OLAP_PIPE(BSX, UX)                                                                                  \
    return eager ? launch_persistent(k_rowwise_pipe<NDIM, GLDS, BSX, UX, true>, a, num_cus, dyn_lds, a.ntiles, s, BSX) \
                 : launch_persistent(k_rowwise_pipe<NDIM, GLDS, BSX, UX, false>, a, num_cus, dyn_lds, a.ntiles, s, BSX)
        if (sh.bs == 1024) OLAP_PIPE(1024, 2);
        OLAP_PIPE(512, 2);
#undef OLAP_PIPE
    }
    a.launched = VEC && NDIM > 1 && a.eager1 ? OLAP_PIPE_STREAM2 : 0;
    if (VEC && NDIM > 1 && a.eager1) {
        // step 1's keys streamed with step 0's
        switch (shape) {
        case 0: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 512, 4, AGG_DIRECT, true>, a, num_cus, dyn_lds, a.ntiles, s, 512);
        case 1: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 512, 2, AGG_DIRECT, true>, a, num_cus, dyn_lds, a.ntiles, s, 512);
        case 2: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 1024, 2, AGG_DIRECT, true>, a, num_cus, dyn_lds, a.ntiles, s, 1024);
        case 3: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 1024, 4, AGG_DIRECT, true>, a, num_cus, dyn_lds, a.ntiles, s, 1024);
        case 4: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 256, 8, AGG_DIRECT, true>, a, num_cus, dyn_lds, a.ntiles, s, 256);
        default: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 512, 8, AGG_DIRECT, true>, a, num_cus, dyn_lds, a.ntiles, s, 512);
        }
    }
    switch (shape) {
    case 0: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 512, 4>, a, num_cus, dyn_lds, a.ntiles, s, 512);
    case 1: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 512, 2>, a, num_cus, dyn_lds, a.ntiles, s, 512);
    case 2: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 1024, 2>, a, num_cus, dyn_lds, a.ntiles, s, 1024);
    case 3: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 1024, 4>, a, num_cus, dyn_lds, a.ntiles, s, 1024);
    case 4: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 256, 8>, a, num_cus, dyn_lds, a.ntiles, s, 256);
    default: return launch_persistent(k_rowwise<NDIM, VEC, GLDS, 512, 8>, a, num_cus, dyn_lds, a.ntiles, s, 512);
    }
}

hipError_t run_rowwise(ScanArgs &a, int ndims, int num_cus, uint32_t dyn_lds, bool vec,
                       int shape, hipStream_t s)
{
    if (shape < 0 || shape >= kNumRowShapes) shape = 0;
    a.launched = 0;
    if (!vec) {
        // unaligned columns: scalar loads, default shape only
        a.ntiles = (a.rows + kTileRW - 1) / kTileRW;
        if (a.ntiles == 0) return hipSuccess;
        if (a.gt_lds) {
            switch (ndims) {
            case 1: return launch_persistent(k_rowwise<1, false, true, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
            case 2: return launch_persistent(k_rowwise<2, false, true, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
            case 3: return launch_persistent(k_rowwise<3, false, true, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
            default: return launch_persistent(k_rowwise<4, false, true, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
            }
        }
        switch (ndims) {
        case 1: return launch_persistent(k_rowwise<1, false, false, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
        case 2: return launch_persistent(k_rowwise<2, false, false, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
        case 3: return launch_persistent(k_rowwise<3, false, false, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
        default: return launch_persistent(k_rowwise<4, false, false, kBlock, kUnroll>, a, num_cus, dyn_lds, a.ntiles, s);
        }
    }
    if (a.gt_lds) {
        switch (ndims) {
        case 1: return rowwise_shape<1, true, true>(a, num_cus, dyn_lds, shape, s);
        case 2: return rowwise_shape<2, true, true>(a, num_cus, dyn_lds, shape, s);
        case 3: return rowwise_shape<3, true, true>(a, num_cus, dyn_lds, shape, s);
        default: return rowwise_shape<4, true, true>(a, num_cus, dyn_lds, shape, s);
        }
    }
    switch (ndims) {
    case 1: return rowwise_shape<1, true, false>(a, num_cus, dyn_lds, shape, s);
    case 2: return rowwise_shape<2, true, false>(a, num_cus, dyn_lds, shape, s);
    case 3: return rowwise_shape<3, true, false>(a, num_cus, dyn_lds, shape, s);
    default: return rowwise_shape<4, true, false>(a, num_cus, dyn_lds, shape, s);
    }
}

// Staged row-wise scan: one fixed launch shape and an explicit grid, so that
// the COUNT and the SCATTER pass give every workgroup the same tiles.
constexpr int kStageBS = 512;
constexpr int kStageU = 4;

template <int MODE>
static const void *staged_kernel(int ndims)
{
    switch (ndims) {
    case 1: return (const void *)k_rowwise<1, true, false, kStageBS, kStageU, MODE>;
    case 2: return (const void *)k_rowwise<2, true, false, kStageBS, kStageU, MODE>;
    case 3: return (const void *)k_rowwise<3, true, false, kStageBS, kStageU, MODE>;
    default: return (const void *)k_rowwise<4, true, false, kStageBS, kStageU, MODE>;
    }
}

hipError_t rowwise_staged_grid(int ndims, int num_cus, uint32_t dyn_lds, uint64_t rows,
                               uint32_t *grid)
{
    int best = 1 << 30;
    for (const void *k : {staged_kernel<AGG_COUNT>(ndims), staged_kernel<AGG_SCATTER>(ndims)}) {
        if (dyn_lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds);
            if (e != hipSuccess) return e;
        }
        int per_cu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kStageBS, dyn_lds);
        if (e != hipSuccess) return e;
        best = std::min(best, per_cu);
    }
    if (best < 1) return hipErrorInvalidConfiguration;
    const uint64_t tiles = (rows + (uint64_t)kStageBS * 4u * kStageU - 1) / ((uint64_t)kStageBS * 4u * kStageU);
    *grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)best * num_cus, tiles));
    return hipSuccess;
}

// one-pass staged form: tiles of kChunkBS threads x kChunkU chunks of 4
// rows, one pool chunk (kChunkRows records) per tile and partition
constexpr int kChunkBS = 1024;
constexpr int kChunkU = 4;
constexpr uint32_t kChunkRows = (uint32_t)kChunkBS * 4u * kChunkU;

static const void *chunk_kernel(int ndims)
{
    switch (ndims) {
    case 1: return (const void *)k_rowwise<1, true, false, kChunkBS, kChunkU, AGG_CHUNK>;
    case 2: return (const void *)k_rowwise<2, true, false, kChunkBS, kChunkU, AGG_CHUNK>;
    case 3: return (const void *)k_rowwise<3, true, false, kChunkBS, kChunkU, AGG_CHUNK>;
    default: return (const void *)k_rowwise<4, true, false, kChunkBS, kChunkU, AGG_CHUNK>;
    }
}

uint32_t rowwise_chunk_rows() { return kChunkRows; }
ScanShape chunk_shape() { return ScanShape{kChunkBS, kChunkU}; }
ScanShape staged_shape() { return ScanShape{kStageBS, kStageU}; }

hipError_t rowwise_chunk_grid(int ndims, int num_cus, uint32_t dyn_lds, uint64_t rows, uint32_t *grid)
{
    const void *k = chunk_kernel(ndims);
    if (dyn_lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds);
        if (e != hipSuccess) return e;
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kChunkBS, dyn_lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const uint64_t tiles = (rows + kChunkRows - 1) / kChunkRows;
    *grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)per_cu * num_cus, tiles));
    return hipSuccess;
}

hipError_t run_rowwise_chunk(ScanArgs &a, int ndims, uint32_t dyn_lds, uint32_t grid, hipStream_t s)
{
    a.ntiles = (a.rows + kChunkRows - 1) / kChunkRows;
    if (a.ntiles == 0) return hipSuccess;
    void *args[] = {&a};
    return hipLaunchKernel(chunk_kernel(ndims), dim3(grid), dim3(kChunkBS), args, dyn_lds, s);
}

hipError_t run_rowwise_staged(ScanArgs &a, int ndims, uint32_t dyn_lds, int mode, uint32_t grid,
                              hipStream_t s)
{
    a.ntiles = (a.rows + (uint64_t)kStageBS * 4u * kStageU - 1) / ((uint64_t)kStageBS * 4u * kStageU);
    if (a.ntiles == 0) return hipSuccess;
    const void *k = mode == AGG_COUNT ? staged_kernel<AGG_COUNT>(ndims) : staged_kernel<AGG_SCATTER>(ndims);
    void *args[] = {&a};
    return hipLaunchKernel(k, dim3(grid), dim3(kStageBS), args, dyn_lds, s);
}

int num_row_shapes() { return kNumRowShapes; }
ScanShape row_shape(int i) { return kRowShapes[i < 0 || i >= kNumRowShapes ? 0 : i]; }

hipError_t run_vwm_sv(ScanArgs &a, int ndims, int num_cus, uint32_t dyn_lds, bool vec,
                      hipStream_t s)
{
    a.ntiles = (a.rows + kTileRW - 1) / kTileRW;
    if (a.ntiles == 0) return hipSuccess;
    OLAP_DISPATCH_ALL(k_vwm_sv)
}

hipError_t run_vwm_dv(ScanArgs &a, int ndims, int num_cus, uint32_t dyn_lds, bool vec,
                      hipStream_t s)
{
    a.ntiles = (a.rows + kTileV - 1) / kTileV;
    if (a.ntiles == 0) return hipSuccess;
    OLAP_DISPATCH_ALL(k_vwm_dv)
}

hipError_t run_cwm_sv_join(ScanArgs &a, int num_cus, uint32_t dyn_lds, bool vec, hipStream_t s)
{
    const uint64_t units = (a.rows + 4ull * kBlock - 1) / (4ull * kBlock);
    if (units == 0) return hipSuccess;
    if (vec) return launch_persistent(k_cwm_sv_join<true>, a, num_cus, dyn_lds, units, s);
    return launch_persistent(k_cwm_sv_join<false>, a, num_cus, dyn_lds, units, s);
}

hipError_t run_cwm_sv_agg(ScanArgs &a, int num_cus, uint32_t dyn_lds, bool vec, hipStream_t s)
{
    const uint64_t units = (a.rows + 4ull * kBlock - 1) / (4ull * kBlock);
    if (units == 0) return hipSuccess;
    if (vec) {
        if (a.gt_lds) return launch_persistent(k_cwm_sv_agg<true, true>, a, num_cus, dyn_lds, units, s);
        return launch_persistent(k_cwm_sv_agg<true, false>, a, num_cus, dyn_lds, units, s);
    }
    if (a.gt_lds) return launch_persistent(k_cwm_sv_agg<false, true>, a, num_cus, dyn_lds, units, s);
    return launch_persistent(k_cwm_sv_agg<false, false>, a, num_cus, dyn_lds, units, s);
}

hipError_t run_cwm_dv_join(ScanArgs &a, int num_cus, uint32_t dyn_lds, bool vec, hipStream_t s)
{
    a.ntiles = (a.rows + kTileV - 1) / kTileV;
    if (a.ntiles == 0) return hipSuccess;
    if (a.step == 0) {
        if (vec) return launch_persistent(k_cwm_dv_join0<true>, a, num_cus, dyn_lds, a.ntiles, s);
        return launch_persistent(k_cwm_dv_join0<false>, a, num_cus, dyn_lds, a.ntiles, s);
    }
    return launch_persistent(k_cwm_dv_join, a, num_cus, dyn_lds, a.ntiles, s);
}

hipError_t run_cwm_dv_agg(ScanArgs &a, int num_cus, uint32_t dyn_lds, hipStream_t s)
{
    a.ntiles = (a.rows + kTileV - 1) / kTileV;
    if (a.ntiles == 0) return hipSuccess;
    if (a.gt_lds) return launch_persistent(k_cwm_dv_agg<true>, a, num_cus, dyn_lds, a.ntiles, s);
    return launch_persistent(k_cwm_dv_agg<false>, a, num_cus, dyn_lds, a.ntiles, s);
}

// ------------------------------------------------------------------------
// sub-step operators
// ------------------------------------------------------------------------

constexpr int kVgBS = 1024;
constexpr int kVgU = 4;  // 128-bit chunks per lane per loop pass

// every row of vector_group: op(k, m1, m2) with k = key - 1 in [0, G); a key
// outside that range is counted in the return value. VEC: 16-byte aligned
// columns, 4 keys (and measures) per 128-bit load. NEED_M: the measures are
// read (the count pass of the staged form reads only the keys). The rows a
// workgroup visits depend only on the grid, so two passes with one grid
// visit the same rows per workgroup.
template <bool VEC, bool NEED_M, typename OP>
__device__ __forceinline__ uint32_t vg_each(int64_t rows, const int32_t *key1, const int32_t *m1,
                                            const int32_t *m2, int64_t G, OP op)
{
    uint32_t nbad = 0;
    auto one = [&](int32_t key, int32_t a, int32_t b) {
        const int64_t k = (int64_t)key - 1;
        if (k < 0 || k >= G) {
            nbad++;
            return;
        }
        op((uint32_t)k, a, b);
    };
    int64_t done = 0;
    if (VEC) {
        const int64_t chunks = rows / 4;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        const v4i *k4 = (const v4i *)key1;
        const v4i *a4 = (const v4i *)m1;
        const v4i *b4 = (const v4i *)m2;
        for (int64_t c0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c0 < chunks; c0 += stride * kVgU) {
            v4i kk[kVgU], ma[kVgU], mb[kVgU];
#pragma unroll
            for (int u = 0; u < kVgU; u++) {
                const int64_t c = c0 + u * stride;
                const bool in = c < chunks;
                kk[u] = in ? __builtin_nontemporal_load(&k4[c]) : v4i{0, 0, 0, 0};
                ma[u] = NEED_M && in ? __builtin_nontemporal_load(&a4[c]) : v4i{0, 0, 0, 0};
                mb[u] = NEED_M && in && m2 ? __builtin_nontemporal_load(&b4[c]) : v4i{0, 0, 0, 0};
            }
#pragma unroll
            for (int u = 0; u < kVgU; u++) {
                if (c0 + u * stride >= chunks) break;
#pragma unroll
                for (int j = 0; j < 4; j++) one(kk[u][j], ma[u][j], mb[u][j]);
            }
        }
        done = chunks * 4;
    }
    for (int64_t i = done + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * blockDim.x)
        one(key1[i], NEED_M ? m1[i] : 0, NEED_M && m2 ? m2[i] : 0);
    return nbad;
}

__device__ __forceinline__ void vg_bad(uint32_t nbad, uint32_t *bad)
{
    for (int o = 32; o > 0; o >>= 1) nbad += __shfl_xor(nbad, o);
    if ((threadIdx.x & 63) == 0 && nbad) atomic_add_u32_dev(bad, nbad);
}

// group vector in LDS (G <= kVgLdsSlots), flushed with one atomic per slot;
// or, beyond LDS and where the staged form does not fit, global atomics
template <bool VEC, bool GLDS>
__global__ __launch_bounds__(kVgBS) void k_vector_group(int64_t rows, const int32_t *key1,
                                                        const int32_t *m1, const int32_t *m2,
                                                        unsigned long long *gv, int64_t G,
                                                        uint32_t *bad)
{
    extern __shared__ __align__(16) uint32_t lds[];
    typedef __attribute__((address_space(3))) unsigned long long l_u64;
    l_u64 *lv = (l_u64 *)lds;
    if (GLDS) {
        for (int64_t g = threadIdx.x; g < G; g += kVgBS) lv[g] = 0ull;
        __syncthreads();
    }
    const uint32_t nbad = vg_each<VEC, true>(rows, key1, m1, m2, G, [&](uint32_t k, int32_t a, int32_t b) {
        const unsigned long long v = (unsigned long long)((long long)a + (long long)b);
        if (GLDS) __hip_atomic_fetch_add(&lv[k], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else atomic_add_u64_dev(&gv[k], v);
    });
    vg_bad(nbad, bad);
    if (GLDS) {
        __syncthreads();
        for (int64_t g = threadIdx.x; g < G; g += kVgBS)
            if (lv[g]) atomic_add_u64_dev(&gv[g], lv[g]);
    }
}

// staged form beyond LDS: the COUNT pass counts rows per (workgroup,
// partition of kPartSlots keys), the SCATTER pass writes one record {key
// within the partition, m1[, m2]} per row into the workgroup's exact range;
// the reduce (olap_stage.hip) adds them up per partition in LDS
template <bool VEC, int MODE>
__global__ __launch_bounds__(kVgBS) void k_vector_group_staged(int64_t rows, const int32_t *key1,
                                                               const int32_t *m1, const int32_t *m2,
                                                               int64_t G, uint32_t *bad, StageArgs st)
{
    extern __shared__ __align__(16) uint32_t lds[];
    stage_begin<MODE>(st, lds);
    __syncthreads();
    const uint32_t nbad = vg_each<VEC, MODE == AGG_SCATTER>(
        rows, key1, m1, m2, G, [&](uint32_t k, int32_t a, int32_t b) {
            if (MODE == AGG_COUNT) stage_count(st, lds, k >> kPartBits);
            else stage_scatter(st, lds, k >> kPartBits, k & (kPartSlots - 1u), a, b);
        });
    if (MODE == AGG_COUNT) vg_bad(nbad, bad);
    __syncthreads();
    stage_end<MODE>(st, lds);
}

static bool vg_vec(const int32_t *key1, const int32_t *m1, const int32_t *m2)
{
    auto al = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    return al(key1) && al(m1) && (!m2 || al(m2));
}

// grid of a vector_group kernel: workgroups per CU from occupancy (with
// lds_bytes of dynamic LDS) on every CU, no more than the rows fill
template <typename K>
static hipError_t vg_grid(K kernel, uint32_t lds_bytes, int num_cus, int64_t rows, uint32_t *grid)
{
    if (lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kVgBS, lds_bytes);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    *grid = (uint32_t)grid_for(rows, kVgBS * 4 * kVgU, per_cu * num_cus);
    return hipSuccess;
}

hipError_t launch_vector_group(int64_t rows, const int32_t *key1, const int32_t *m1,
                               const int32_t *m2, unsigned long long *gv, int64_t G,
                               unsigned long long *total, uint32_t *bad, int num_cus,
                               hipStream_t s)
{
    (void)total;
    if (rows <= 0) return hipSuccess;
    const bool vec = vg_vec(key1, m1, m2);
    const bool glds = G <= (int64_t)kVgLdsSlots;
    const uint32_t lds = glds ? (uint32_t)(G * 8) : 0u;
    uint32_t grid = 1;
    hipError_t e;
#define OLAP_VG(V, L)                                                                          \
    e = vg_grid(k_vector_group<V, L>, lds, num_cus, rows, &grid);                              \
    if (e != hipSuccess) return e;                                                             \
    k_vector_group<V, L><<<grid, kVgBS, lds, s>>>(rows, key1, m1, m2, gv, G, bad);
    if (vec && glds) { OLAP_VG(true, true) }
    else if (vec) { OLAP_VG(true, false) }
    else if (glds) { OLAP_VG(false, true) }
    else { OLAP_VG(false, false) }
#undef OLAP_VG
    return hipGetLastError();
}

hipError_t vector_group_staged_grid(int64_t rows, const int32_t *key1, const int32_t *m1, const int32_t *m2,
                                    uint32_t lds_bytes, int num_cus, uint32_t *grid)
{
    uint32_t g0 = 1, g1 = 1;
    hipError_t e;
    if (vg_vec(key1, m1, m2)) {
        e = vg_grid(k_vector_group_staged<true, AGG_COUNT>, lds_bytes, num_cus, rows, &g0);
        if (e == hipSuccess) e = vg_grid(k_vector_group_staged<true, AGG_SCATTER>, lds_bytes, num_cus, rows, &g1);
    } else {
        e = vg_grid(k_vector_group_staged<false, AGG_COUNT>, lds_bytes, num_cus, rows, &g0);
        if (e == hipSuccess) e = vg_grid(k_vector_group_staged<false, AGG_SCATTER>, lds_bytes, num_cus, rows, &g1);
    }
    *grid = std::min(g0, g1);
    return e;
}

hipError_t run_vector_group_staged(int mode, int64_t rows, const int32_t *key1, const int32_t *m1,
                                   const int32_t *m2, int64_t G, uint32_t *bad, const StageArgs &st,
                                   uint32_t lds_bytes, uint32_t grid, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    const bool vec = vg_vec(key1, m1, m2);
    if (mode == AGG_COUNT) {
        if (vec) k_vector_group_staged<true, AGG_COUNT><<<grid, kVgBS, lds_bytes, s>>>(rows, key1, m1, m2, G, bad, st);
        else k_vector_group_staged<false, AGG_COUNT><<<grid, kVgBS, lds_bytes, s>>>(rows, key1, m1, m2, G, bad, st);
    } else {
        if (vec) k_vector_group_staged<true, AGG_SCATTER><<<grid, kVgBS, lds_bytes, s>>>(rows, key1, m1, m2, G, bad, st);
        else k_vector_group_staged<false, AGG_SCATTER><<<grid, kVgBS, lds_bytes, s>>>(rows, key1, m1, m2, G, bad, st);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_truncate_u32(const unsigned long long *v, int64_t n,
                                                      uint32_t *out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)v[i];
}

hipError_t launch_truncate_u32(const unsigned long long *v, int64_t n, uint32_t *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_truncate_u32<<<grid_for(n, 256 * 4, 4096), 256, 0, s>>>(v, n, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_sum_u64(const unsigned long long *v, int64_t n,
                                                 unsigned long long *out)
{
    unsigned long long acc = 0ull;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        acc += v[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0 && acc) atomic_add_u64_dev(out, acc);
}

hipError_t launch_sum_u64(const unsigned long long *v, int64_t n, unsigned long long *out,
                          int num_cus, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_sum_u64<<<grid_for(n, 256 * 8, num_cus * 4), 256, 0, s>>>(v, n, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_agg_expr(int64_t rows, const int32_t *p,
                                                  const int32_t *q, const int32_t *t,
                                                  unsigned long long *out)
{
    long long acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * blockDim.x) {
        const long long pq = (long long)p[i] * (long long)q[i];
        acc += pq * (long long)t[i] + 100ll * pq;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0 && acc) atomic_add_u64_dev(out, (unsigned long long)acc);
}

hipError_t launch_agg_expr(int64_t rows, const int32_t *p, const int32_t *q, const int32_t *t,
                           unsigned long long *out, int num_cus, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    k_agg_expr<<<grid_for(rows, 256 * 8, num_cus * 4), 256, 0, s>>>(rows, p, q, t, out);
    return hipGetLastError();
}

}  // namespace olap
