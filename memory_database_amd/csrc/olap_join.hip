// olap_join.hip -- the dense-vector join sub-step on gfx950: build a vector
// indexed by key from a build relation, probe it with a probe relation's
// keys and sum what the probes find. It is the star join's probe step on its
// own: the vector sits in LDS when it fits (up to kJoinLdsEntries entries,
// copied once per workgroup) and is gathered through L2 otherwise; the probe
// keys stream through 128-bit loads.
#include <algorithm>

#include "olap_internal.h"

namespace olap {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kJoinBS = 512;
constexpr int64_t kJoinLdsEntries = kLdsBytes / 4;  // the whole LDS (160 KiB)

__device__ __forceinline__ void add_u64(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add_u32(uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

int grid_of(int64_t items, int per_block, int cap)
{
    int64_t g = (items + per_block - 1) / per_block;
    return (int)(g < 1 ? 1 : g > cap ? cap : g);
}

// Build, pass 0: every key that a row names starts unclaimed (-1).
__global__ __launch_bounds__(256) void k_join_unclaim(int64_t rows, const int32_t *key,
                                                      int32_t *winner, int64_t vlen)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = key[i];
        if (k >= 0 && k < vlen) winner[k] = -1;
    }
}

// Build, pass 1: the last row naming a key claims it (largest row index), as
// the reference's sequential loop leaves the last row's payload.
__global__ __launch_bounds__(256) void k_join_claim(int64_t rows, const int32_t *key,
                                                    int32_t *winner, int64_t vlen, uint32_t *bad)
{
    uint32_t nbad = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = key[i];
        if (k < 0 || k >= vlen) {
            nbad++;
            continue;
        }
        __hip_atomic_fetch_max(&winner[k], (int32_t)i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (int o = 32; o > 0; o >>= 1) nbad += __shfl_xor(nbad, o);
    if ((threadIdx.x & 63) == 0 && nbad) add_u32(bad, nbad);
}

// Build, pass 2: the claiming rows write their payloads.
__global__ __launch_bounds__(256) void k_join_write(int64_t rows, const int32_t *key,
                                                    const int32_t *payload, const int32_t *winner,
                                                    int32_t *vec, int64_t vlen)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = key[i];
        if (k >= 0 && k < vlen && winner[k] == (int32_t)i) vec[k] = payload[i];
    }
}

// Probe: sum of vec[key] over the probe keys, in int64. LDS = the vector is
// copied into LDS first. VEC = keys are 16-byte aligned: 4 keys per lane per
// 128-bit load, 4 loads in flight per lane.
template <bool LDS, bool VEC, int BS = kJoinBS>
__global__ __launch_bounds__(BS) void k_join_probe(int64_t rows, const int32_t *key,
                                                   const int32_t *vec, int64_t vlen,
                                                   unsigned long long *total, uint32_t *bad)
{
    extern __shared__ __align__(16) int32_t lvec[];
    if (LDS) {
        for (int64_t e = threadIdx.x; e < vlen; e += BS) lvec[e] = vec[e];
        __syncthreads();
    }
    auto look = [&](int32_t k, long long &acc, uint32_t &nbad) {
        if ((uint32_t)k >= (uint64_t)vlen) {
            nbad++;
            return;
        }
        if (LDS) {
            const __attribute__((address_space(3))) int32_t *l =
                (const __attribute__((address_space(3))) int32_t *)lvec;
            acc += l[k];
        } else {
            acc += vec[k];
        }
    };
    long long acc = 0;
    uint32_t nbad = 0;
    int64_t done = 0;
    if (VEC) {
        constexpr int U = 4;
        const int64_t chunks = rows / 4;
        const v4i *k4 = (const v4i *)key;
        const int64_t stride = (int64_t)gridDim.x * BS;
        for (int64_t c0 = blockIdx.x * (int64_t)BS + threadIdx.x; c0 < chunks; c0 += stride * U) {
            v4i kk[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t c = c0 + u * stride;
                kk[u] = c < chunks ? __builtin_nontemporal_load(&k4[c]) : v4i{-1, -1, -1, -1};
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (c0 + u * stride >= chunks) break;
                look(kk[u].x, acc, nbad);
                look(kk[u].y, acc, nbad);
                look(kk[u].z, acc, nbad);
                look(kk[u].w, acc, nbad);
            }
        }
        done = chunks * 4;
    }
    for (int64_t i = done + blockIdx.x * (int64_t)BS + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * BS)
        look(key[i], acc, nbad);
    for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o);
        nbad += __shfl_xor(nbad, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (acc) add_u64(total, (unsigned long long)acc);
        if (nbad) add_u32(bad, nbad);
    }
}

}  // namespace

hipError_t launch_join_build(int64_t rows, const int32_t *key, const int32_t *payload,
                             int32_t *winner, int32_t *vec, int64_t vlen, uint32_t *bad,
                             int num_cus, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    const int grid = grid_of(rows, 256 * 8, num_cus * 8);
    k_join_unclaim<<<grid, 256, 0, s>>>(rows, key, winner, vlen);
    k_join_claim<<<grid, 256, 0, s>>>(rows, key, winner, vlen, bad);
    k_join_write<<<grid, 256, 0, s>>>(rows, key, payload, winner, vec, vlen);
    return hipGetLastError();
}

hipError_t launch_join_probe(int64_t rows, const int32_t *key, const int32_t *vec, int64_t vlen,
                             unsigned long long *total, uint32_t *bad, int num_cus, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    const bool vec4 = ((uintptr_t)key & 15u) == 0;
    if (vlen <= kJoinLdsEntries) {
        // one workgroup per CU pass copies the vector once; 2 per CU overlap
        // one's copy with the other's probes. A vector above 80 KiB leaves
        // room for one workgroup per CU, which then takes 1024 threads.
        const uint32_t lds = (uint32_t)std::max<int64_t>(vlen, 4) * 4u;
        const bool big = lds > 80 * 1024;
        const void *k = big ? (vec4 ? (const void *)k_join_probe<true, true, 1024> : (const void *)k_join_probe<true, false, 1024>)
                            : (vec4 ? (const void *)k_join_probe<true, true> : (const void *)k_join_probe<true, false>);
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        const int bs = big ? 1024 : kJoinBS;
        const int grid = grid_of(rows, bs * 64, num_cus * (big ? 1 : 2));
        void *args[] = {&rows, &key, &vec, &vlen, &total, &bad};
        hipError_t e = hipLaunchKernel(k, dim3(grid), dim3(bs), args, lds, s);
        if (e != hipSuccess) return e;
    } else {
        const int grid = grid_of(rows, kJoinBS * 16, num_cus * 8);
        if (vec4) k_join_probe<false, true><<<grid, kJoinBS, 0, s>>>(rows, key, vec, vlen, total, bad);
        else k_join_probe<false, false><<<grid, kJoinBS, 0, s>>>(rows, key, vec, vlen, total, bad);
    }
    return hipGetLastError();
}

}  // namespace olap
