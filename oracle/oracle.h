/* CPU oracle for the cusort-hip sorting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for bench.py, as the timed CPU baseline).  The product path never
 * calls it.
 *
 * It restates the reference's correctness oracle:
 *   - expected output = host std::sort of every array / segment
 *     (reference bench.cu:300-301, test.cu:46-48);
 *   - record checks of validateBatch (reference bench.cu:209-228): the key
 *     sequence equals the sorted reference, every payload lies in range, stays
 *     in its own array, names an input record equal to the output record, and
 *     appears exactly once;
 *   - the reference's stateless input hash mixBits (bench.cu:152-155) and its
 *     integer distributions (bench.cu:161-191, bench_distributions.hpp:11-19);
 *   - the added distributions of the MI355X configurations (uniform full
 *     range, 16 distinct values, bounded perturbation, organ pipe, alternating
 *     extremes, 50% duplicates), fp32 IEEE-754 totalOrder keys, uint64 keys,
 *     int64+int32 records and ragged CSR segments, which have no reference
 *     counterpart and use the same std::sort oracle per segment.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_splitmix64    published SplitMix64 outputs (seed 0 and 1234567)
 *   orc_mixbits       composes to the identity with the published inverse
 *                     hash (lowbias32_r)
 *   orc_generate      closed forms of the structured distributions, value
 *                     ranges, committed golden fixtures
 *   orc_sort_*        numpy.sort (an independent library routine), brute
 *                     force over all permutations on tiny inputs, closed-form
 *                     sorted outputs, committed golden fixtures
 *   orc_validate_*    planted mistakes (swapped key, lost / duplicated record,
 *                     record moved across arrays, payload / key mismatch)
 *                     must be reported; correct outputs must pass
 *   totalOrder        IEEE-754 totalOrder special cases (-NaN < -inf < -0 <
 *                     +0 < +inf < +NaN, payload order of NaNs)
 */
#ifndef CUSORT_ORACLE_H
#define CUSORT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types; numbering matches include/cusort_hip.h (cs_dtype). */
enum {
    ORC_I32 = 0,        /* int32 key */
    ORC_REC_I32_I32 = 1,/* {int32 key, int32 payload}, 8 B, layout of IntAndData */
    ORC_F32 = 2,        /* fp32 key, IEEE-754 totalOrder */
    ORC_U64 = 3,        /* uint64 key */
    ORC_REC_I64_I32 = 4,/* {int64 key, int32 payload, int32 pad}, 16 B */
    ORC_U32 = 5,        /* uint32 key */
    ORC_I64 = 6,        /* int64 key */
    ORC_F64 = 7         /* fp64 key, IEEE-754 totalOrder */
};

/* Distributions.  0-8 are the nine seeded distributions of the batched and
 * record configurations; 9-11 restate the reference's EXTREMES, RANDOM and
 * BINARY cases (bench.cu:171, 183-187); 12 is the fp32 N(0,1)+specials case;
 * 13 is the 10%-presorted-run case of the large-array configuration.        */
enum {
    ORC_D_UNIFORM = 0,
    ORC_D_SORTED = 1,
    ORC_D_REVERSED = 2,
    ORC_D_CONSTANT = 3,
    ORC_D_DISTINCT16 = 4,
    ORC_D_PERTURBED = 5,
    ORC_D_ORGAN_PIPE = 6,
    ORC_D_ALT_EXTREMES = 7,
    ORC_D_DUP50 = 8,
    ORC_D_EXTREMES6 = 9,
    ORC_D_RANDOM1M = 10,
    ORC_D_BINARY = 11,
    ORC_D_NORMAL_SPECIAL = 12,
    ORC_D_PRESORTED_RUNS = 13,
    ORC_D_COUNT = 14
};

uint64_t orc_splitmix64(uint64_t x);
uint32_t orc_mixbits(uint32_t x);
uint32_t orc_f32_to_ordered(uint32_t bits);
uint32_t orc_ordered_to_f32(uint32_t key);

/* Fill `out` with `count` elements of `dtype`; element g belongs to array
 * g / n at local position g % n (n = array length; for one large array pass
 * n = count).  Records get payload = g.  Returns 0 or -1 on bad arguments. */
int orc_generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out);

/* Per-array / per-segment std::sort, keys compared by the dtype's order
 * (records by key only).  `threads` <= 0 uses all host cores.              */
int orc_sort_batched(int dtype, const void* in, void* out, int64_t n, int64_t arrays, int threads);
int orc_sort_segmented(int dtype, const void* in, void* out, const int64_t* offsets,
                       int64_t segments, int threads);
/* Whole-array parallel sort (__gnu_parallel::sort over `threads` cores).   */
int orc_sort_large(int dtype, const void* in, void* out, int64_t count, int threads);

/* Validators.  Return 0 when `result` is a correct sort of `input`, else the
 * 1-based index of the first bad element (or -1 for argument errors); a
 * human-readable reason is written to msg.                                  */
int64_t orc_validate_batched(int dtype, const void* input, const void* result, int64_t n,
                             int64_t arrays, char* msg, int msglen);
int64_t orc_validate_segmented(int dtype, const void* input, const void* result,
                               const int64_t* offsets, int64_t segments, char* msg, int msglen);

/* CPU baseline: median wall seconds over `trials` of the per-array
 * (or per-segment, or whole-array) std::sort on `threads` host threads,
 * restoring the input before every trial; *threads_used reports the count. */
double orc_time_batched(int dtype, const void* in, int64_t n, int64_t arrays, int threads,
                        int trials, int* threads_used);
double orc_time_segmented(int dtype, const void* in, const int64_t* offsets, int64_t segments,
                          int threads, int trials, int* threads_used);
double orc_time_large(int dtype, const void* in, int64_t count, int threads, int trials,
                      int* threads_used);
int orc_host_cores(void);

#ifdef __cplusplus
}
#endif

#endif
