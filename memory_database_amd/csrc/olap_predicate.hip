// olap_predicate.hip -- general dimension predicates on gfx950: a program of
// up to OLAP_MAX_TERMS range / IN / hop terms, combined left to right, is
// evaluated for every dimension row in one pass and written as a dimension
// vector (the code of a passing row, NULL otherwise). Dimension tables are
// small (at most a few million rows), so one grid-stride pass with coalesced
// column reads is enough; IN lists are binary-searched in a sorted device
// copy and hop keys gather from the parent vector through L2.
#include "olap_internal.h"

namespace olap {

namespace {

__device__ __forceinline__ bool in_sorted(const int32_t *v, int32_t n, int32_t x)
{
    int32_t lo = 0, hi = n;  // first index with v[i] >= x
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (v[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && v[lo] == x;
}

__global__ __launch_bounds__(256) void k_dim_program(DevProgram p, int64_t rows, int32_t code_bytes,
                                                     void *out, uint32_t *err)
{
    uint32_t nerr = 0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        bool acc = false;
        bool gfound = false;
        int32_t gcode = 0;
        for (int i = 0; i < p.nterms; i++) {
            const DevTerm &t = p.term[i];
            bool ok;
            if (t.kind == OLAP_TERM_RANGE) {
                const int32_t v = t.col[r];
                ok = v >= t.lo && v < t.hi;
            } else if (t.kind == OLAP_TERM_IN) {
                ok = in_sorted(t.values, t.nvalues, t.col[r]);
            } else {
                const int64_t k = (int64_t)t.col[r] + t.key_bias;
                int32_t code = -1;
                if (k >= 0 && k < t.parent_rows) {
                    if (t.parent_code_bytes == 1) {
                        const int8_t c = ((const int8_t *)t.parent)[k];
                        code = c == (int8_t)OLAP_DIM_NULL_I8 ? -1 : (int32_t)c;
                    } else {
                        code = ((const int32_t *)t.parent)[k];
                    }
                }
                ok = code != -1;
                if (i == p.group_term) {
                    gfound = ok;
                    gcode = code;
                }
            }
            if (t.negate) ok = !ok;
            acc = i == 0 ? ok : t.combine == OLAP_OR ? (acc || ok) : (acc && ok);
        }
        int32_t code = 0;
        if (p.group == OLAP_GROUP_COLUMN) {
            const int32_t g = p.group_col[r];
            code = p.group_mod > 0 ? (int32_t)((uint32_t)g % (uint32_t)p.group_mod) : g;
        } else if (p.group == OLAP_GROUP_HOP) {
            acc = acc && gfound;
            code = gcode;
        }
        if (code_bytes == 1) {
            if (acc && (code < 0 || code >= OLAP_DIM_NULL_I8)) {
                nerr++;
                acc = false;
            }
            ((int8_t *)out)[r] = acc ? (int8_t)code : (int8_t)OLAP_DIM_NULL_I8;
        } else {
            ((int32_t *)out)[r] = acc ? code : OLAP_DIM_NULL_I32;
        }
    }
    for (int o = 32; o > 0; o >>= 1) nerr += __shfl_xor(nerr, o);
    if ((threadIdx.x & 63) == 0 && nerr)
        __hip_atomic_fetch_add(err, nerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

hipError_t launch_dim_program(const DevProgram &p, int64_t rows, int32_t code_bytes, void *out,
                              uint32_t *err, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    int64_t g = (rows + 255) / 256;
    if (g > 4096) g = 4096;
    k_dim_program<<<(int)g, 256, 0, s>>>(p, rows, code_bytes, out, err);
    return hipGetLastError();
}

}  // namespace olap
