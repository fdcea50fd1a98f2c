// Shared definitions of the eggman HIP library: status handling and element-type traits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <stdarg.h>
#include <string>

#include "../../include/eggman_hip.h"

namespace eggman {

// Sets the thread-local error message and returns `code`.
int fail(int code, const char *fmt, ...);

#define EGG_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::eggman::fail(e_ == hipErrorOutOfMemory ? EGGMAN_ENOMEM : EGGMAN_EHIP, \
                                  "%s failed: %s", #call, hipGetErrorString(e_));          \
    } while (0)

#define EGG_TRY(expr)            \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != EGGMAN_OK)    \
            return rc_;          \
    } while (0)

// ------------------------------------------------------------------ element types on device
//
// Every muladd below is written as a separate multiply and add: the sources are compiled with
// fp-contract off, so no fused multiply-add is formed and a state's sum matches the reference's
// _dp_states! (src/dp.jl:344-363) bit for bit when accumulated in plan order.  Complex multiply
// is (ac - bd, ad + bc), Julia's `*` for Complex.  The product is rounded before it meets the
// accumulator, so muladd_hh(acc, a, b) == add_h(acc, mul_hh(a, b)) and muladd(acc, p, h) ==
// add_h(acc, mul_ph(p, h)) bit for bit: products may be formed in any order as long as they are
// added in the list's.

// The tail of a gather loop taken in batches of U: `batch(n)` with n an integral_constant equal to
// rem (0 < rem < U), so the last rem loads are issued together like a full batch's instead of one
// round trip each.
template <int U, class F>
__device__ inline void batch_tail(int rem, F &&batch) {
    if constexpr (U > 1) {
        if (rem == U - 1) batch(std::integral_constant<int, U - 1>{});
        else batch_tail<U - 1>(rem, batch);
    }
}

template <int DT> struct Elt;

template <> struct Elt<EGGMAN_F64> {
    using H = double;   // state value
    using P = double;   // pair value
    __device__ static inline H zero() { return 0.0; }
    __device__ static inline H one() { return 1.0; }
    __device__ static inline H muladd(H acc, P p, H h) { return acc + p * h; }
    __device__ static inline H muladd_hh(H acc, H a, H b) { return acc + a * b; }
    __device__ static inline H mul_hh(H a, H b) { return a * b; }
    __device__ static inline H mul_ph(P p, H h) { return p * h; }
    __device__ static inline H add_h(H a, H b) { return a + b; }
};
template <> struct Elt<EGGMAN_F32> {
    using H = float;
    using P = float;
    __device__ static inline H zero() { return 0.0f; }
    __device__ static inline H one() { return 1.0f; }
    __device__ static inline H muladd(H acc, P p, H h) { return acc + p * h; }
    __device__ static inline H muladd_hh(H acc, H a, H b) { return acc + a * b; }
    __device__ static inline H mul_hh(H a, H b) { return a * b; }
    __device__ static inline H mul_ph(P p, H h) { return p * h; }
    __device__ static inline H add_h(H a, H b) { return a + b; }
};
template <> struct Elt<EGGMAN_C64> {
    using H = double2;
    using P = double2;
    __device__ static inline H zero() { return make_double2(0.0, 0.0); }
    __device__ static inline H one() { return make_double2(1.0, 0.0); }
    __device__ static inline H muladd(H acc, P p, H h) {
        const double re = p.x * h.x - p.y * h.y;
        const double im = p.x * h.y + p.y * h.x;
        return make_double2(acc.x + re, acc.y + im);
    }
    __device__ static inline H muladd_hh(H acc, H a, H b) { return muladd(acc, a, b); }
    __device__ static inline H mul_hh(H a, H b) {
        return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
    __device__ static inline H mul_ph(P p, H h) { return mul_hh(p, h); }
    __device__ static inline H add_h(H a, H b) { return make_double2(a.x + b.x, a.y + b.y); }
};
template <> struct Elt<EGGMAN_C32> {
    using H = float2;
    using P = float2;
    __device__ static inline H zero() { return make_float2(0.0f, 0.0f); }
    __device__ static inline H one() { return make_float2(1.0f, 0.0f); }
    __device__ static inline H muladd(H acc, P p, H h) {
        const float re = p.x * h.x - p.y * h.y;
        const float im = p.x * h.y + p.y * h.x;
        return make_float2(acc.x + re, acc.y + im);
    }
    __device__ static inline H muladd_hh(H acc, H a, H b) { return muladd(acc, a, b); }
    __device__ static inline H mul_hh(H a, H b) {
        return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
    __device__ static inline H mul_ph(P p, H h) { return mul_hh(p, h); }
    __device__ static inline H add_h(H a, H b) { return make_float2(a.x + b.x, a.y + b.y); }
};
// Exact integer walk: states are 128-bit integers modulo 2^128, pair values are int64.  The
// DP only adds and multiplies, so the result is exact whenever the true hafnian fits in 127
// bits, however large the intermediate values get.
template <> struct Elt<EGGMAN_I128> {
    using H = ulonglong2;   // (x = lo, y = hi)
    using P = long long;
    __device__ static inline H zero() { return make_ulonglong2(0ull, 0ull); }
    __device__ static inline H one() { return make_ulonglong2(1ull, 0ull); }
    __device__ static inline H muladd(H acc, P p, H h) {
        const unsigned long long pl = (unsigned long long)p;
        const unsigned long long ph = p < 0 ? ~0ull : 0ull;     // sign extension
        return add(acc, mul(make_ulonglong2(pl, ph), h));
    }
    // a * b modulo 2^128
    __device__ static inline H mul(H a, H b) {
        const unsigned long long lo = a.x * b.x;
        const unsigned long long hi = __umul64hi(a.x, b.x) + a.x * b.y + a.y * b.x;
        return make_ulonglong2(lo, hi);
    }
    __device__ static inline H add(H a, H b) {
        const unsigned long long lo = a.x + b.x;
        return make_ulonglong2(lo, a.y + b.y + (lo < a.x ? 1ull : 0ull));
    }
    __device__ static inline H muladd_hh(H acc, H a, H b) { return add(acc, mul(a, b)); }
    __device__ static inline H mul_hh(H a, H b) { return mul(a, b); }
    __device__ static inline H mul_ph(P p, H h) {
        return mul(make_ulonglong2((unsigned long long)p, p < 0 ? ~0ull : 0ull), h);
    }
    __device__ static inline H add_h(H a, H b) { return add(a, b); }
};

inline int elt_hbytes(int dtype) {
    switch (dtype) {
    case EGGMAN_F32: return 4;
    case EGGMAN_F64: return 8;
    case EGGMAN_C32: return 8;
    case EGGMAN_C64: return 16;
    case EGGMAN_I128: return 16;
    }
    return 0;
}
inline int elt_pbytes(int dtype) {
    switch (dtype) {
    case EGGMAN_F32: return 4;
    case EGGMAN_F64: return 8;
    case EGGMAN_C32: return 8;
    case EGGMAN_C64: return 16;
    case EGGMAN_I128: return 8;
    }
    return 0;
}

}  // namespace eggman
