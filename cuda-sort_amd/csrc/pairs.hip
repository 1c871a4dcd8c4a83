// Key / payload pairs held as two arrays (structure of arrays) through the
// ragged and large-array sorts: the pairs are packed into the library's
// records ({int32 key, int32 payload} = the reference's IntAndData,
// intanddata.hpp:9-79, or {int64 key, int32 payload, pad}), sorted by the
// record path with its tie rule, and unpacked again.  The result is exactly
// the record sort of {key[i], payload[i]} split apart.  The batched sort
// does the same without the detour (its tile loads the two arrays directly,
// batched_pairs.hip).
#include <hip/hip_runtime.h>

#include "../../include/cusort_hip.h"
#include "elem.hpp"
#include "gen.hpp"

namespace cs {
namespace {

template <class R, class K>
__global__ void k_pack_pairs(const K* keys, const int32_t* vals, R* rec, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        R r{};
        r.key = keys[i];
        r.payload = vals[i];
        rec[i] = r;
    }
}

template <class R, class K>
__global__ void k_unpack_pairs(const R* rec, K* keys, int32_t* vals, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const R r = rec[i];
        keys[i] = r.key;
        vals[i] = r.payload;
    }
}

unsigned grid_for(int64_t n) {
    const int64_t g = (n + 255) / 256;
    return unsigned(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

}  // namespace

int pair_record_dtype(int key_dtype) {
    return key_dtype == CS_I32 ? CS_REC_I32_I32 : (key_dtype == CS_I64 ? CS_REC_I64_I32 : -1);
}

hipError_t pairs_pack(int key_dtype, const void* keys, const int32_t* vals, void* rec, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (key_dtype == CS_I32)
        hipLaunchKernelGGL((k_pack_pairs<RecI32, int32_t>), dim3(grid_for(n)), dim3(256), 0, s,
                           static_cast<const int32_t*>(keys), vals, static_cast<RecI32*>(rec), n);
    else
        hipLaunchKernelGGL((k_pack_pairs<RecI64, int64_t>), dim3(grid_for(n)), dim3(256), 0, s,
                           static_cast<const int64_t*>(keys), vals, static_cast<RecI64*>(rec), n);
    return hipGetLastError();
}

hipError_t pairs_unpack(int key_dtype, const void* rec, void* keys, int32_t* vals, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (key_dtype == CS_I32)
        hipLaunchKernelGGL((k_unpack_pairs<RecI32, int32_t>), dim3(grid_for(n)), dim3(256), 0, s,
                           static_cast<const RecI32*>(rec), static_cast<int32_t*>(keys), vals, n);
    else
        hipLaunchKernelGGL((k_unpack_pairs<RecI64, int64_t>), dim3(grid_for(n)), dim3(256), 0, s,
                           static_cast<const RecI64*>(rec), static_cast<int64_t*>(keys), vals, n);
    return hipGetLastError();
}

}  // namespace cs

using namespace cs;

extern "C" int cs_sort_pairs(int key_dtype, void* keys, int32_t* vals, int64_t n, cs_workspace* ws,
                             hipStream_t stream) {
    const int rd = pair_record_dtype(key_dtype);
    if (rd < 0 || n < 0 || (n > 0 && (!keys || !vals))) return CS_E_INVALID_VALUE;
    if (n == 0) return CS_OK;
    void* rec = nullptr;
    hipError_t e = pool_malloc_async(&rec, size_t(n) * cs_dtype_size(rd), stream);
    if (e != hipSuccess) return int(e);
    e = pairs_pack(key_dtype, keys, vals, rec, n, stream);
    int rc = e == hipSuccess ? cs_sort_hybrid(rd, rec, n, ws, stream) : int(e);
    if (rc == CS_OK) {
        e = pairs_unpack(key_dtype, rec, keys, vals, n, stream);
        if (e != hipSuccess) rc = int(e);
    }
    const hipError_t f = hipFreeAsync(rec, stream);
    return rc == CS_OK && f != hipSuccess ? int(f) : rc;
}
