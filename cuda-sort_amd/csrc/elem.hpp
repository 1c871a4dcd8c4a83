// Element types of the sorting path: storage format in HBM, the key that is
// compared, the payload that travels with it, and the padding element.
//
// Reference semantics (cusort.hpp:257-287, intanddata.hpp:9-79): arrays are
// padded to a power of two with the maximum element, records compare by key
// only, and a compare-exchange keeps each side's own element on ties
// (cusort.hpp:460-472).  fp32 keys are sorted through the IEEE-754 totalOrder
// bit transform so that NaN, -0.0 and infinities have a total order and no key
// arithmetic is ever performed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cusort_hip.h"

namespace cs {

struct RecI32 { int32_t key; int32_t payload; };                 // = reference IntAndData
struct alignas(16) RecI64 { int64_t key; int32_t payload; int32_t pad; };

template <int DT> struct Elem;

template <> struct Elem<CS_I32> {
    using Storage = int32_t; using Key = int32_t;
    static constexpr bool kPayload = false;
    static constexpr Key kPad = INT32_MAX;
    __device__ static Key key(Storage s) { return s; }
    __device__ static int32_t payload(Storage) { return 0; }
    __device__ static Storage make(Key k, int32_t) { return k; }
};
template <> struct Elem<CS_U32> {
    using Storage = uint32_t; using Key = uint32_t;
    static constexpr bool kPayload = false;
    static constexpr Key kPad = 0xFFFFFFFFu;
    __device__ static Key key(Storage s) { return s; }
    __device__ static int32_t payload(Storage) { return 0; }
    __device__ static Storage make(Key k, int32_t) { return k; }
};
// fp32 stored as raw IEEE bits; key = totalOrder transform (sign bit set ->
// flip all bits, else flip the sign bit), an order-preserving bijection onto
// uint32.  The padding key 0xFFFFFFFF is never written back.
template <> struct Elem<CS_F32> {
    using Storage = uint32_t; using Key = uint32_t;
    static constexpr bool kPayload = false;
    static constexpr Key kPad = 0xFFFFFFFFu;
    __device__ static Key key(Storage b) { return b ^ (uint32_t(int32_t(b) >> 31) | 0x80000000u); }
    __device__ static int32_t payload(Storage) { return 0; }
    __device__ static Storage make(Key k, int32_t) { return k ^ (uint32_t(int32_t(~k) >> 31) | 0x80000000u); }
};
// fp64 likewise: raw IEEE bits, totalOrder key over uint64.
template <> struct Elem<CS_F64> {
    using Storage = uint64_t; using Key = uint64_t;
    static constexpr bool kPayload = false;
    static constexpr Key kPad = ~uint64_t(0);
    __device__ static Key key(Storage b) { return b ^ (uint64_t(int64_t(b) >> 63) | (uint64_t(1) << 63)); }
    __device__ static int32_t payload(Storage) { return 0; }
    __device__ static Storage make(Key k, int32_t) { return k ^ (uint64_t(int64_t(~k) >> 63) | (uint64_t(1) << 63)); }
};
template <> struct Elem<CS_U64> {
    using Storage = uint64_t; using Key = uint64_t;
    static constexpr bool kPayload = false;
    static constexpr Key kPad = ~uint64_t(0);
    __device__ static Key key(Storage s) { return s; }
    __device__ static int32_t payload(Storage) { return 0; }
    __device__ static Storage make(Key k, int32_t) { return k; }
};
template <> struct Elem<CS_I64> {
    using Storage = int64_t; using Key = int64_t;
    static constexpr bool kPayload = false;
    static constexpr Key kPad = INT64_MAX;
    __device__ static Key key(Storage s) { return s; }
    __device__ static int32_t payload(Storage) { return 0; }
    __device__ static Storage make(Key k, int32_t) { return k; }
};
template <> struct Elem<CS_REC_I32_I32> {
    using Storage = RecI32; using Key = int32_t;
    static constexpr bool kPayload = true;
    static constexpr Key kPad = INT32_MAX;        // IntAndData::max() = {INT_MAX, INT_MAX}
    __device__ static Key key(Storage s) { return s.key; }
    __device__ static int32_t payload(Storage s) { return s.payload; }
    __device__ static Storage make(Key k, int32_t p) { return RecI32{k, p}; }
};
template <> struct Elem<CS_REC_I64_I32> {
    using Storage = RecI64; using Key = int64_t;
    static constexpr bool kPayload = true;
    static constexpr Key kPad = INT64_MAX;
    __device__ static Key key(Storage s) { return s.key; }
    __device__ static int32_t payload(Storage s) { return s.payload; }
    __device__ static Storage make(Key k, int32_t p) { return RecI64{k, p, 0}; }
};

// Compare-exchange of the element pair at stored positions (lo, hi).
// Ascending: the lower position receives the smaller key.  An exchange
// happens only on a strict inversion, so on equal keys each side keeps its
// own element (cusort.hpp:460-472): no record is lost or duplicated, and
// maximum-key padding never moves to a lower position.
template <class E, bool kDesc>
__device__ __forceinline__ void cmpx(typename E::Key& ka, int32_t& pa, typename E::Key& kb, int32_t& pb) {
    using K = typename E::Key;
    if constexpr (!E::kPayload && sizeof(K) == 4) {
        const K lo = ka < kb ? ka : kb;   // v_min
        const K hi = ka < kb ? kb : ka;   // v_max
        ka = kDesc ? hi : lo;
        kb = kDesc ? lo : hi;
    } else {
        const bool swap = kDesc ? (ka < kb) : (kb < ka);
        const K k0 = swap ? kb : ka;
        const K k1 = swap ? ka : kb;
        ka = k0; kb = k1;
        if constexpr (E::kPayload) {
            const int32_t p0 = swap ? pb : pa;
            const int32_t p1 = swap ? pa : pb;
            pa = p0; pb = p1;
        }
    }
}

}  // namespace cs
