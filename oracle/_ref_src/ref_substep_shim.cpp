// Recipe piece for oracle/_ref: the reference's star-join, vector-join and
// group-by sub-step operators, compiled from their own source files where
// they lie.
// Two translation units would each define the header's non-inline helpers, so
// this shim is built twice, once per operator file (SUBSTEP selects it).
#if SUBSTEP == 1
#define main ref_starjoin_main
#include "code/starjoin/test_starjoin_option.cpp"
#undef main
extern "C" int ref_starjoin_rowwise(int n, const int8_t *c, const int8_t *s, const int8_t *p,
                                    const int8_t *d, const int32_t *fc, const int32_t *fs,
                                    const int32_t *fp, const int32_t *fd)
{
    return starjoinalgo_rowwise(n, c, s, p, d, fc, fs, fp, fd);
}
// result must be zeroed by the caller (the reference adds into it)
extern "C" int ref_starjoin_cwm_sv(int n, const int8_t *c, const int8_t *s, const int8_t *p,
                                   const int8_t *d, const int32_t *fc, const int32_t *fs,
                                   const int32_t *fp, const int32_t *fd, int *result)
{
    return starjoinalgo_cwm_sv(n, c, s, p, d, fc, fs, fp, fd, result);
}
#elif SUBSTEP == 3
#define main ref_join_main
#include "code/join/test_join_option.cpp"
#undef main
// vecjoin build (:30-36) and probe (:58-64) on a relation of n rows
extern "C" void ref_join_build(long long n, int32_t *key, int32_t *payload, int *vector)
{
    relation_t r;
    r.key = key;
    r.payload = payload;
    r.num_tuples = (uint64_t)n;
    build(&r, vector);
}
extern "C" int ref_join_probe(long long n, int32_t *key, int *vector, int prefetch)
{
    relation_t s;
    s.key = key;
    s.payload = nullptr;
    s.num_tuples = (uint64_t)n;
    return prefetch ? probe_prefetch(&s, vector) : probe(&s, vector);
}
#else
#define main ref_group_main
#include "code/group/test_group_option.cpp"
#undef main
extern "C" long long ref_vector_group(long long n, const int *key1, const int *m1, const int *m2,
                                      int *group_vector, int group_num)
{
    size_t nn = (size_t)n;
    return (long long)vector_group(nn, key1, m1, m2, group_vector, group_num);
}
#endif
