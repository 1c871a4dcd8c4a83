// Batched sort of key / payload pairs held as two arrays (structure of
// arrays): the {int32 key, int32 payload} and {int64 key, int32 payload}
// records of the batched sort, with keys and payloads in separate buffers.
// The register tile is the records' own (keys and payloads already live in
// separate registers and LDS regions), so the network, its tie rule and the
// result are exactly those of the record sort; only global I/O differs:
// lane-striped element loads and stores of each array, keys and payloads
// apart, 12 (8) bytes per element each way instead of the 16 (8) of the
// padded record.
#include <hip/hip_runtime.h>

#include "../../include/cusort_hip.h"
#include "batched.hpp"

namespace cs {
namespace {

// One array of `valid` pairs (valid <= W; 0 for an idle group) with the T
// lanes t of this group; slot r of lane t is array position r * T + t on
// both sides of the network (lane-striped), re-laid to and from the identity
// layout through LDS as the record tile does.
template <class E, int LOGW, int LOGE>
__device__ __forceinline__ void tile_sort_pairs(const typename E::Key* ks, const int32_t* ps, typename E::Key* kd,
                                                int32_t* pd, int valid, int t, typename E::Key* lk, int32_t* lp) {
    using C = TileCfg<E, LOGW, LOGE>;
    constexpr int NE = C::NE;
    constexpr int T = C::T;
    constexpr int kBufLast = C::kMultiWave ? (C::NTRANS & 1) * C::APB * C::REGION : 0;
    constexpr int kBufFirst = C::kMultiWave ? C::APB * C::REGION : 0;
    Tile<E, LOGW, LOGE> tile;
    const bool full = (NE - 1) * T + t < valid;   // every slot of this lane in range
    if (full) {
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            tile.k[r] = ks[r * T + t];
            tile.p[r] = ps[r * T + t];
        });
    } else {
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            const int g = r * T + t;
            tile.k[r] = g < valid ? ks[g] : E::kPad;
            tile.p[r] = g < valid ? ps[g] : INT32_MAX;
        });
    }
    if constexpr (C::NTRANS > 0) tile.template io_transpose<1, true>(t, lk, lp, kBufFirst);
    tile.template run<0>(t, lk, lp);
    if constexpr (C::NTRANS > 0) tile.template io_transpose<1, false>(t, lk, lp, kBufLast);
    if (full) {
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            kd[r * T + t] = tile.k[r];
            pd[r * T + t] = tile.p[r];
        });
    } else {
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            const int g = r * T + t;
            if (g < valid) {
                kd[g] = tile.k[r];
                pd[g] = tile.p[r];
            }
        });
    }
}

template <class E, int LOGW, int LOGE>
__global__ void __launch_bounds__((TileCfg<E, LOGW, LOGE>::BLOCK))
k_bitonic_tile_pairs(const typename E::Key* ks, const int32_t* ps, typename E::Key* kd, int32_t* pd, int64_t arrays,
                     int n) {
    using C = TileCfg<E, LOGW, LOGE>;
    static_assert(E::kPayload, "pairs carry a payload");
    __shared__ typename E::Key lds_k[C::LDS_KEYS];
    __shared__ int32_t lds_p[C::LDS_PAYLOAD];
    const int tid = int(threadIdx.x);
    const int group = tid >> C::LOGT;
    const int t = tid & (C::T - 1);
    const int64_t a = int64_t(blockIdx.x) * C::APB + group;
    const bool active = a < arrays;
    const int64_t base = (active ? a : 0) * int64_t(n);
    tile_sort_pairs<E, LOGW, LOGE>(ks + base, ps + base, kd + base, pd + base, active ? n : 0, t,
                                   lds_k + (C::NTRANS ? group * C::REGION : 0),
                                   lds_p + ((C::NTRANS && E::kPayload) ? group * C::REGION : 0));
}

template <int DT, int LW>
hipError_t launch_pairs(const void* ks, const int32_t* ps, void* kd, int32_t* pd, int64_t arrays, int n,
                        hipStream_t s) {
    using E = Elem<DT>;
    using K = typename E::Key;
    constexpr int LE = auto_loge(DT, LW);
    using C = TileCfg<E, LW, LE>;
    const int64_t blocks = (arrays + C::APB - 1) / C::APB;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_bitonic_tile_pairs<E, LW, LE>), dim3(unsigned(blocks)), dim3(C::BLOCK), 0, s,
                       static_cast<const K*>(ks), ps, static_cast<K*>(kd), pd, arrays, n);
    return hipGetLastError();
}

using PairsFn = hipError_t (*)(const void*, const int32_t*, void*, int32_t*, int64_t, int, hipStream_t);

template <int DT>
PairsFn pairs_kernel(int logw) {
    PairsFn f = nullptr;
    hfor<0, max_logw(DT) + 1>([&](auto w_) {
        constexpr int LW = decltype(w_)::value;
        if (LW == logw) f = &launch_pairs<DT, LW>;
    });
    return f;
}

}  // namespace
}  // namespace cs

extern "C" int cs_batched_sort_pairs(int key_dtype, const void* keys_in, const int32_t* vals_in, void* keys_out,
                                     int32_t* vals_out, int64_t arrays, int32_t n, hipStream_t stream) {
    using namespace cs;
    if (key_dtype != CS_I32 && key_dtype != CS_I64) return CS_E_INVALID_VALUE;
    if (arrays < 0 || n < 0) return CS_E_INVALID_VALUE;
    const int rec = key_dtype == CS_I32 ? CS_REC_I32_I32 : CS_REC_I64_I32;
    if (n > cs_batched_max_n(rec)) return CS_E_CAPACITY;
    if (arrays == 0 || n == 0) return CS_OK;
    if (!keys_in || !vals_in || !keys_out || !vals_out) return CS_E_INVALID_VALUE;
    if (arrays > (int64_t(1) << 40)) return CS_E_INVALID_VALUE;
    // in place or disjoint, per buffer pair
    const size_t count = size_t(arrays) * size_t(n);
    auto partial = [](const void* a, const void* b, size_t bytes) {
        const char* x = static_cast<const char*>(a);
        const char* y = static_cast<const char*>(b);
        return x != y && x < y + bytes && y < x + bytes;
    };
    if (partial(keys_in, keys_out, count * (key_dtype == CS_I32 ? 4 : 8)) || partial(vals_in, vals_out, count * 4))
        return CS_E_INVALID_VALUE;
    int logw = 0;
    while ((1 << logw) < n) ++logw;
    const PairsFn f = rec == CS_REC_I32_I32 ? pairs_kernel<CS_REC_I32_I32>(logw) : pairs_kernel<CS_REC_I64_I32>(logw);
    if (!f) return CS_E_UNSUPPORTED;
    const hipError_t e = f(keys_in, vals_in, keys_out, vals_out, arrays, n, stream);
    return e == hipSuccess ? CS_OK : int(e);
}
