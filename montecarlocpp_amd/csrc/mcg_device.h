// mcg_device.h -- device-side building blocks of the trajectory engine.
//
// Counter-based random numbers (Philox4x32-10, Salmon et al., SC'11), the
// integer alias draw, the reference's direction samplers (random.cpp:16-44)
// and precision traits for the FP64 and FP32-flight paths.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "mcg_math.h"

namespace mcg {
namespace dev {

// ------------------------------------------------------------------ Philox
// a ^ b ^ k in one VALU op (gfx950 V_BITOP3_B32, truth table 0x96); the
// round key k is wavefront-uniform and sits in an SGPR
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t k) {
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(k));
    return r;
}

__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32 x 32 -> 64 multiply (v_mad_u64_u32) gives both halves
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0, p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t lo0 = uint32_t(p0), hi0 = uint32_t(p0 >> 32);
        const uint32_t lo1 = uint32_t(p1), hi1 = uint32_t(p1 >> 32);
        const uint32_t n0 = xor3(hi1, c1, k0), n2 = xor3(hi0, c3, k1);
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

// One event's stream: block j of event s of bundle n under key (k0, k1).
struct Event {
    uint32_t k0, k1, s, nlo, nhi;
    __device__ __forceinline__ uint4 block(uint32_t j) const { return philox(j, s, nlo, nhi, k0, k1); }
};

__device__ __forceinline__ uint32_t word_of(const uint4& b, int i) {
    return i == 0 ? b.x : i == 1 ? b.y : i == 2 ? b.z : b.w;
}

// Walker alias draw from one 32-bit word: slot = floor(x K / 2^32), the low
// 32 bits of x K are the acceptance fraction.  Table entry = thr | alias<<32.
__device__ __forceinline__ uint32_t alias_draw(const uint64_t* table, uint32_t n, uint32_t x) {
    const uint64_t prod = uint64_t(x) * uint64_t(n);
    const uint32_t slot = uint32_t(prod >> 32);
    const uint64_t e = table[slot];
    return uint32_t(prod) < uint32_t(e) ? slot : uint32_t(e >> 32);
}

// ------------------------------------------------------- precision traits
template <typename F>
struct Real;

template <>
struct Real<double> {
    static constexpr double kMin = DBL_MIN;
    // 53-bit uniform on [0, 1): 27 bits of the high word a, 26 of the low b
    __device__ static __forceinline__ double uniform(uint32_t a, uint32_t b) {
        return (double(a >> 5) * 67108864.0 + double(b >> 6)) * (1.0 / 9007199254740992.0);
    }
    __device__ static __forceinline__ double sqrt_(double x) { return sqrt(x); }
    // sin(pi r), cos(pi r) and log from the portable header, bit-identical
    // to the CPU checker's evaluation
    __device__ static __forceinline__ void sincos_pi(double r, double& s, double& c) {
        mcg_math::sincospi(r, &s, &c);
    }
    __device__ static __forceinline__ double neglog1m(double u) { return -mcg_math::log(1.0 - u); }
    __device__ static __forceinline__ double recip_norm_scale(double x, double n2) {
        return x / sqrt(n2);
    }
    // a / b, correctly rounded (the checker divides)
    __device__ static __forceinline__ double div(double a, double b) { return a / b; }
};

template <>
struct Real<float> {
    static constexpr float kMin = FLT_MIN;
    // 24-bit uniform on [0, 1) from the high word alone (the FP64 uniform's
    // leading bits, so both precisions take the same decisions)
    __device__ static __forceinline__ float uniform(uint32_t a, uint32_t) {
        return float(a >> 8) * 5.9604644775390625e-8f;
    }
    // FP32 flight is compared statistically, so it uses the hardware's
    // one-instruction approximations: v_sqrt_f32, v_rcp_f32 (1 ulp) and
    // v_sin_f32 / v_cos_f32 (argument in turns: sin(pi r) = sin(2 pi (r/2)),
    // r/2 in [-1/2, 1/2), absolute error near 1e-6)
    __device__ static __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static __forceinline__ void sincos_pi(float r, float& s, float& c) {
        s = __builtin_amdgcn_sinf(0.5f * r);
        c = __builtin_amdgcn_cosf(0.5f * r);
    }
    __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    // -ln(1 - u) = -log2(1 - u) ln 2
    __device__ static __forceinline__ float neglog1m(float u) {
        return -__builtin_amdgcn_logf(1.0f - u) * 0.69314718055994531f;
    }
    __device__ static __forceinline__ float recip_norm_scale(float x, float n2) {
        return x * __builtin_amdgcn_rsqf(n2);
    }
};

template <typename F>
struct V3 {
    F x, y, z;
};

// FP64 keeps the checker's separately rounded products and sums; FP32 flight
// fuses them (v_fma_f32).
template <typename F>
__device__ __forceinline__ F dot3(const F* n, const V3<F>& v) {
    if constexpr (sizeof(F) == 4)
        return __builtin_fmaf(n[2], v.z, __builtin_fmaf(n[1], v.y, n[0] * v.x));
    else
        return n[0] * v.x + n[1] * v.y + n[2] * v.z;
}

// column-major 3x3 times vector, row by row as (m0 x + m3 y) + m6 z
template <typename F>
__device__ __forceinline__ V3<F> matvec(const F* m, const V3<F>& v) {
    if constexpr (sizeof(F) == 4) {
        return V3<F>{__builtin_fmaf(m[6], v.z, __builtin_fmaf(m[3], v.y, m[0] * v.x)),
                     __builtin_fmaf(m[7], v.z, __builtin_fmaf(m[4], v.y, m[1] * v.x)),
                     __builtin_fmaf(m[8], v.z, __builtin_fmaf(m[5], v.y, m[2] * v.x))};
    } else {
        return V3<F>{m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
                     m[2] * v.x + m[5] * v.y + m[8] * v.z};
    }
}

// Row r of matvec(m, v), evaluated exactly as matvec evaluates it.
template <typename F>
__device__ __forceinline__ F matrow(const F* m, int r, const V3<F>& v) {
    if constexpr (sizeof(F) == 4)
        return __builtin_fmaf(m[r + 6], v.z, __builtin_fmaf(m[r + 3], v.y, m[r] * v.x));
    else
        return m[r] * v.x + m[r + 3] * v.y + m[r + 6] * v.z;
}

// Phonon::dir normalises (phonon.cpp:88-93 with Eigen's normalized())
template <typename F>
__device__ __forceinline__ V3<F> normalized(const V3<F>& v) {
    const F n2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (n2 > F(0)) {
        return V3<F>{Real<F>::recip_norm_scale(v.x, n2), Real<F>::recip_norm_scale(v.y, n2),
                     Real<F>::recip_norm_scale(v.z, n2)};
    }
    return v;
}

// drawIso (random.cpp:16-27): cos(theta) ~ U[-1, 1), phi ~ pi U[-1, 1).
// An event's uniforms take their high words from block 0 (h) and their low
// words from block 1 (l), which FP32 never evaluates.
template <typename F>
__device__ __forceinline__ V3<F> draw_iso(const uint4& h, const uint4& l) {
    const F cosT = Real<F>::uniform(h.x, l.x) * F(2) + F(-1);
    const F sinT = Real<F>::sqrt_(F(1) - cosT * cosT);
    const F r = Real<F>::uniform(h.y, l.y) * F(2) + F(-1);
    F s, c;
    Real<F>::sincos_pi(r, s, c);
    return V3<F>{sinT * c, sinT * s, cosT};
}

// drawAniso (random.cpp:29-44): sin^2(theta) = |r|, r ~ U[-1, 1); with
// bidir the sign of r picks the hemisphere
template <typename F>
__device__ __forceinline__ V3<F> draw_aniso(const uint4& h, const uint4& l, bool bidir) {
    const F r = Real<F>::uniform(h.x, l.x) * F(2) + F(-1);
    const F sg = (bidir && r < F(0)) ? F(-1) : F(1);
    const F sin2 = r < F(0) ? -r : r;
    const F sinT = Real<F>::sqrt_(sin2);
    const F cosT = sg * Real<F>::sqrt_(F(1) - sin2);
    const F q = Real<F>::uniform(h.y, l.y) * F(2) + F(-1);
    F s, c;
    Real<F>::sincos_pi(q, s, c);
    return V3<F>{sinT * c, sinT * s, cosT};
}

}  // namespace dev
}  // namespace mcg
