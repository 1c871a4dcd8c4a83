// Batched-sort kernel registry: which (dtype, W, E) tile kernels exist and
// how to launch them.  Each dtype's kernels are instantiated in their own
// translation unit (batched_<dtype>.hip) so they compile in parallel.
#pragma once

#include "bitonic_lds.hpp"
#include "bitonic_tile.hpp"
#include "bitonic_xlane.hpp"

namespace cs {

using LaunchFn = hipError_t (*)(const void* in, void* out, int64_t arrays, int n, int logw, unsigned flags,
                                int threads, hipStream_t stream);

struct KernelInfo {
    LaunchFn launch = nullptr;
    const void* kernel = nullptr;  // device function, for hipFuncGetAttributes
    int block = 0;                 // threads per workgroup (0: chosen at launch)
    int items = 0;                 // elements per lane
    int apb = 0;                   // arrays per workgroup
    int lds = 0;                   // static LDS bytes
};

constexpr int kMaxLogW4 = 14;   // 16384 4-byte keys
constexpr int kMaxLogW8 = 13;   // 8192 8-byte keys / IntAndData records
constexpr int kMaxLogW16 = 12;  // 4096 16-byte records

constexpr int max_logw(int dtype) {
    return (dtype == CS_I32 || dtype == CS_U32 || dtype == CS_F32) ? kMaxLogW4
           : (dtype == CS_REC_I64_I32) ? kMaxLogW16 : kMaxLogW8;
}

// Default elements per lane of the register tile for W = 2^logw: 32 VGPRs
// of keys (16 for 16-byte records).  With E = 32 a wave holds two arrays of
// 1024, an array needs sum_{k>5} ceil(k/5) LDS transposes (10 at W = 1024,
// against 14 at E = 16), and every transpose of W <= 2048 is free of LDS
// bank conflicts under the x + (x >> 5) padding (32 banks served per
// 32-lane half of the wave).
//
// 8-byte elements from W = 4096 up take 16 per lane: with 32, a workgroup's
// two double-buffered arrays need 139 KB of LDS and only one workgroup fits
// a CU (measured, 4M elements: {int32, int32} records at W = 4096 56.9 ->
// 49.2 us, at W = 8192 64.0 -> 61.6 us).  16-byte records always take 16.
constexpr int auto_loge(int dtype, int logw) {
    const bool wide8 = dtype == CS_U64 || dtype == CS_I64 || dtype == CS_F64 || dtype == CS_REC_I32_I32;
    const int cap = (dtype == CS_REC_I64_I32 || (wide8 && logw >= 12)) ? 4 : 5;
    return logw < cap ? logw : cap;
}

template <class E>
constexpr int tile_lds_bytes(int logw, int loge) {
    const int T = 1 << (logw - loge);
    const int block = T >= 256 ? T : 256;
    const int apb = block / T;
    const int ntrans = make_plan(logw, loge).ntrans;
    const int padshift = sizeof(typename E::Key) == 4 ? 5 : 4;
    const int region = (1 << logw) + ((1 << logw) >> padshift);
    const int nbuf = T > 64 ? 2 : 1;
    return ntrans ? nbuf * apb * region * (int(sizeof(typename E::Key)) + (E::kPayload ? 4 : 0)) : 0;
}

template <int DT>
constexpr bool tile_exists(int logw, int loge) {
    if (loge < 0 || loge > logw || logw > max_logw(DT)) return false;
    if (logw > 0 && loge == 0) return false;                    // needs at least one slot bit
    if (logw - loge > 10) return false;                        // at most 1024 lanes
    if (tile_lds_bytes<Elem<DT>>(logw, loge) > 160 * 1024) return false;
    // The default tiling, plus every v6 tiling (items 1..16) for the
    // reference's own element types.
    if (loge == auto_loge(DT, logw)) return true;
    return (DT == CS_I32 || DT == CS_REC_I32_I32) && loge <= 4;
}

// Key-only 4-byte element types sort 32..2048-element arrays with the
// cross-lane register engine (bitonic_xlane.hpp) instead of the LDS-transpose
// tile of the same shape (32 keys per lane, W / 32 lanes per array).  Stages
// from kXlaneWindow on move through LDS windows: at W = 1024 that keeps lane
// positions 4 and 5 (permlane swaps, three issue slots per element) out of
// the network and leaves four LDS passes instead of the tile's ten.
constexpr int kXlaneWindow = 9;
constexpr bool xlane_dtype(int dt) { return dt == CS_I32 || dt == CS_U32 || dt == CS_F32; }
constexpr bool xlane_shape(int dt, int logw, int loge) {
    return xlane_dtype(dt) && loge == xl::kLogE && logw >= xl::kLogE && logw <= 11;
}

template <int DT, int LW>
hipError_t launch_xlane(const void* in, void* out, int64_t arrays, int n, int, unsigned flags, int, hipStream_t s) {
    using S = typename Elem<DT>::Storage;
    constexpr int apb = 4 * (64 >> (LW - xl::kLogE));
    const int64_t blocks = (arrays + apb - 1) / apb;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_bitonic_xlane<Elem<DT>, LW, kXlaneWindow>), dim3(unsigned(blocks)), dim3(256), 0, s,
                       static_cast<const S*>(in), static_cast<S*>(out), arrays, n, flags);
    return hipGetLastError();
}

template <int DT, int LW, int LE>
hipError_t launch_tile(const void* in, void* out, int64_t arrays, int n, int, unsigned flags, int, hipStream_t s) {
    using C = TileCfg<Elem<DT>, LW, LE>;
    using S = typename Elem<DT>::Storage;
    const int64_t blocks = (arrays + C::APB - 1) / C::APB;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_bitonic_tile<Elem<DT>, LW, LE>), dim3(unsigned(blocks)), dim3(C::BLOCK), 0, s,
                       static_cast<const S*>(in), static_cast<S*>(out), arrays, n, flags);
    return hipGetLastError();
}

template <int DT>
hipError_t launch_lds(const void* in, void* out, int64_t arrays, int n, int logw, unsigned flags, int threads,
                      hipStream_t s) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    if (arrays <= 0) return hipSuccess;
    const size_t bytes = (size_t(1) << logw) * (sizeof(typename E::Key) + (E::kPayload ? 4 : 0));
    if (bytes > 64 * 1024) {
        // Dynamic LDS beyond 64 KB must be opted into per kernel.
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bitonic_lds<E>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
        if (e != hipSuccess) {
            (void)hipGetLastError();   // do not leave the error pending for the next call
            return e;
        }
    }
    hipLaunchKernelGGL(k_bitonic_lds<E>, dim3(unsigned(arrays)), dim3(unsigned(threads)), bytes, s,
                       static_cast<const S*>(in), static_cast<S*>(out), arrays, n, logw, flags);
    return hipGetLastError();
}

template <int I, int N, class F>
inline void hfor(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        hfor<I + 1, N>(f);
    }
}

// Tile kernels of `DT` with LO <= log2(W) < HI (one translation unit each).
template <int DT, int LO, int HI>
KernelInfo tile_lookup(int logw, int loge) {
    KernelInfo r{};
    constexpr int hi = HI < max_logw(DT) + 1 ? HI : max_logw(DT) + 1;
    hfor<LO, hi>([&](auto w_) {
        constexpr int LW = decltype(w_)::value;
        hfor<0, 6>([&](auto e_) {
            constexpr int LE = decltype(e_)::value;
            if constexpr (xlane_shape(DT, LW, LE)) {
                if (LW == logw && LE == loge) {
                    r.launch = &launch_xlane<DT, LW>;
                    r.kernel = reinterpret_cast<const void*>(&k_bitonic_xlane<Elem<DT>, LW, kXlaneWindow>);
                    r.block = 256;
                    r.items = xl::kE;
                    r.apb = 4 * (64 >> (LW - xl::kLogE));
                    r.lds = r.apb * xl::region_words(LW) * 4;
                }
            } else if constexpr (tile_exists<DT>(LW, LE)) {
                if (LW == logw && LE == loge) {
                    using C = TileCfg<Elem<DT>, LW, LE>;
                    r.launch = &launch_tile<DT, LW, LE>;
                    r.kernel = reinterpret_cast<const void*>(&k_bitonic_tile<Elem<DT>, LW, LE>);
                    r.block = C::BLOCK;
                    r.items = C::NE;
                    r.apb = C::APB;
                    r.lds = C::LDS_BYTES;
                }
            }
        });
    });
    return r;
}

template <int DT>
KernelInfo lds_lookup() {
    KernelInfo r{};
    r.launch = &launch_lds<DT>;
    r.kernel = reinterpret_cast<const void*>(&k_bitonic_lds<Elem<DT>>);
    r.apb = 1;
    return r;
}

// Defined in batched_<dtype>.hip.
KernelInfo tile_kernel(int dtype, int logw, int loge);
KernelInfo lds_kernel(int dtype);

// Split point of the two translation units per element type.
constexpr int kSplitLogW = 10;
#define CS_DECLARE_TILES(name)                  \
    KernelInfo tile_kernel_##name##_lo(int, int); \
    KernelInfo tile_kernel_##name##_hi(int, int);
CS_DECLARE_TILES(i32)
CS_DECLARE_TILES(u32)
CS_DECLARE_TILES(f32)
CS_DECLARE_TILES(u64)
CS_DECLARE_TILES(f64)
CS_DECLARE_TILES(i64)
CS_DECLARE_TILES(rec32)
CS_DECLARE_TILES(rec64)
#undef CS_DECLARE_TILES
KernelInfo lds_kernel_i32();
KernelInfo lds_kernel_u32();
KernelInfo lds_kernel_f32();
KernelInfo lds_kernel_u64();
KernelInfo lds_kernel_f64();
KernelInfo lds_kernel_i64();
KernelInfo lds_kernel_rec32();
KernelInfo lds_kernel_rec64();

}  // namespace cs
