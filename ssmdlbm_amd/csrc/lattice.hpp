// lattice.hpp -- compile-time velocity sets and per-cell collision algebra
// for the gfx950 kernels.
//
// Direction order and weights are those of the reference velocity sets
// (src/velocitySet/D3Q19.cuh:50-131, D3Q27.cuh:50-140, D3Q7.cuh:50-122):
// the boundary kernels address populations by number and opposite
// directions are the pairs (2k-1, 2k), which the collision below exploits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ssmd {

template <int B, int E, class F>
__device__ __forceinline__ void sfor(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

struct Dir {
    int x, y, z;
};

// A running sum whose first term is taken as is: 0 + x must be kept for
// x = -0 in IEEE arithmetic, so starting from 0.f costs one add per sum.
// `any` is known at compile time in the unrolled loops that use it.
struct Sum {
    float v = 0.f;
    bool any = false;
    __device__ __forceinline__ void add(float d) {
        v = any ? v + d : d;
        any = true;
    }
    __device__ __forceinline__ void sub(float d) {
        v = any ? v - d : -d;
        any = true;
    }
};

// Shared table: D3Q19 is the first 19 entries, D3Q7 the first 7.
constexpr Dir kDir[27] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},   {-1, 0, -1}, {0, 1, 1},   {0, -1, -1}, {1, -1, 0},
    {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},  {0, -1, 1},  {1, 1, 1},   {-1, -1, -1},
    {1, 1, -1},  {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};

template <int Q>
constexpr float weight(int i) {
    if constexpr (Q == 19) {
        return i == 0 ? float(1.0 / 3.0) : (i <= 6 ? float(1.0 / 18.0) : float(1.0 / 36.0));
    } else {
        return i == 0 ? float(8.0 / 27.0)
                      : (i <= 6 ? float(2.0 / 27.0) : (i <= 18 ? float(1.0 / 54.0) : float(1.0 / 216.0)));
    }
}
constexpr float weight7(int k) { return k == 0 ? 0.25f : 0.125f; }

constexpr float kCs2 = float(1.0 / 3.0);

// --------------------------------------------------------------------------
// Population storage formats (deviation fbar = f - w is what is stored).
// pair(a, b, one, sum, dif): the FP32 sum and difference of two stored
// values, exactly rounded (`one` is 1.0f held opaque by the caller).
struct PopF32 {
    using T = float;
    static __device__ __forceinline__ float load(T v) { return v; }
    static __device__ __forceinline__ T store(float v) { return v; }
    static __device__ __forceinline__ void pair(T a, T b, float, float &sum, float &dif) {
        sum = a + b;
        dif = a - b;
    }
};
struct PopF16 {
    using T = uint16_t;
    static __device__ __forceinline__ float load(T v) { return float(__builtin_bit_cast(_Float16, v)); }
    // IEEE conversion of the FP32 value in the default round-to-nearest-even
    // mode (the reference's __float2half of its float result, utils.cuh:77-83).
    // The value is made opaque first: otherwise the compiler folds the last
    // FMA that produced it into a mixed-precision FMA with an f16 result
    // (v_fma_mixlo_f16), which rounds once from the exact product-sum and so
    // differs from rounding the FP32 result in rare ties.
    static __device__ __forceinline__ T store(float v) {
        asm volatile("" : "+v"(v));
        return __builtin_bit_cast(uint16_t, _Float16(v));
    }
    // a * 1 + b rounds exactly as a + b; as an FMA of f16 operands it is one
    // mixed-precision instruction (v_fma_mix_f32) instead of two conversions
    // and an add
    static __device__ __forceinline__ void pair(T a, T b, float one, float &sum, float &dif) {
        const float x = load(a), y = load(b);
        sum = __builtin_fmaf(x, one, y);
        dif = __builtin_fmaf(x, one, -y);
    }
};
struct PopBF16 {
    using T = uint16_t;
    static __device__ __forceinline__ float load(T v) { return __builtin_bit_cast(float, uint32_t(v) << 16); }
    // gfx950 converts with round-to-nearest-even in one instruction
    // (v_cvt_pk_bf16_f32).
    static __device__ __forceinline__ T store(float v) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v)); }
    static __device__ __forceinline__ void pair(T a, T b, float, float &sum, float &dif) {
        const float x = load(a), y = load(b);
        sum = x + y;
        dif = x - y;
    }
};

// Two post-collision populations of one cell rounded to a 16-bit storage
// format in one instruction (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32,
// round-to-nearest-even), returned as one word: the first in the low half.
// The values are made opaque first, as P::store does, so that no FMA is
// folded into the rounding.
template <class P> struct Pack2;
template <> struct Pack2<PopF16> {
    typedef _Float16 H2 __attribute__((ext_vector_type(2)));
    typedef float F2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ uint32_t pack(float a, float b) {
        asm volatile("" : "+v"(a), "+v"(b));
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(F2{a, b}, H2));
    }
};
template <> struct Pack2<PopBF16> {
    typedef __bf16 B2 __attribute__((ext_vector_type(2)));
    typedef float F2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ uint32_t pack(float a, float b) {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(F2{a, b}, B2));
    }
};

// --------------------------------------------------------------------------
// Isotropic gradient 3 * sum_i w_i c_i v(x + c_i) over opposite pairs,
// grouped by weight class (reference phaseField.cuh:158-218, lbm.cu:157-217).
// `at(dx,dy,dz)` returns the stencil value.
template <int Q, class At>
__device__ __forceinline__ void iso_grad(At &&at, float &gx, float &gy, float &gz) {
    Sum s[3][3];  // [class][axis]
    sfor<0, (Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * P + 1;
        constexpr Dir c = kDir[i];
        constexpr int cls = i <= 6 ? 0 : (i <= 18 ? 1 : 2);
        const float d = at(c.x, c.y, c.z) - at(-c.x, -c.y, -c.z);
        if constexpr (c.x != 0) { if (c.x > 0) s[cls][0].add(d); else s[cls][0].sub(d); }
        if constexpr (c.y != 0) { if (c.y > 0) s[cls][1].add(d); else s[cls][1].sub(d); }
        if constexpr (c.z != 0) { if (c.z > 0) s[cls][2].add(d); else s[cls][2].sub(d); }
    });
    // 3 w_class, the prefactor folded into the weights
    constexpr float w1 = 3.0f * weight<Q>(1), w2 = 3.0f * weight<Q>(7), w3 = Q == 27 ? 3.0f * weight<Q>(19) : 0.f;
    float r[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = w1 * s[0][a].v + w2 * s[1][a].v;
        if constexpr (Q == 27) v += w3 * s[2][a].v;
        r[a] = v;
    }
    gx = r[0];
    gy = r[1];
    gz = r[2];
}

// --------------------------------------------------------------------------
// Regularized collision of one cell (reference lbm.cu:33-128 moments,
// lbm.cu:226-269 collision; D3Q19.cuh / D3Q27.cuh equilibrium, Hermite
// non-equilibrium and Guo force).
//
// Moments use closed forms of the reference sums, exact in real arithmetic
// for both lattices (they carry 4th-order isotropy):
//   sum_i c_a c_b fbar_eq_i = rho u_a u_b + cs2 (rho-1) delta_ab
//   sum_i c_a c_b F_i       = u_a F_b + u_b F_a
// so Pi = traceless( S - rho u u + (u F + F u)/2 ),  S = sum_i c c fbar_i.
// Post-collision populations are split into even and odd parts in c and
// emitted in opposite pairs.
struct Cell {
    float rho, drho, ux, uy, uz;
    float pxx, pyy, pzz, pxy, pxz, pyz;
};

// The rest state rho = 1, u = 0, Pi = 0: with F = 0 every population and
// phase moment collide() and phase_moments() emit from it is exactly +0.
__device__ __forceinline__ Cell rest_or(bool keep, const Cell &m) {
    Cell r;
    r.rho = keep ? m.rho : 1.0f;
    r.drho = keep ? m.drho : 0.f;
    r.ux = keep ? m.ux : 0.f;
    r.uy = keep ? m.uy : 0.f;
    r.uz = keep ? m.uz : 0.f;
    r.pxx = keep ? m.pxx : 0.f;
    r.pyy = keep ? m.pyy : 0.f;
    r.pzz = keep ? m.pzz : 0.f;
    r.pxy = keep ? m.pxy : 0.f;
    r.pxz = keep ? m.pxz : 0.f;
    r.pyz = keep ? m.pyz : 0.f;
    return r;
}

// `get(i)` returns the stored element of direction i for the cell.
template <int Q, class P, class Get>
__device__ __forceinline__ Cell moments_of(Get &&get, float fx, float fy, float fz) {
    Cell m;
    float one = 1.0f;
    asm volatile("" : "+v"(one));  // not folded: P::pair multiplies by it
    float dr = P::load(get(0));
    Sum jx, jy, jz, sxx, syy, szz, sxy, sxz, syz;
    sfor<0, (Q - 1) / 2>([&](auto Pp) {
        constexpr int i = 2 * Pp + 1;
        constexpr Dir c = kDir[i];
        float sum, dif;
        P::pair(get(i), get(i + 1), one, sum, dif);
        dr += sum;
        auto signed_add = [](Sum &a, int sg, float v) {
            if (sg > 0) a.add(v);
            else a.sub(v);
        };
        if constexpr (c.x != 0) { signed_add(jx, c.x, dif); sxx.add(sum); }
        if constexpr (c.y != 0) { signed_add(jy, c.y, dif); syy.add(sum); }
        if constexpr (c.z != 0) { signed_add(jz, c.z, dif); szz.add(sum); }
        if constexpr (c.x * c.y != 0) signed_add(sxy, c.x * c.y, sum);
        if constexpr (c.x * c.z != 0) signed_add(sxz, c.x * c.z, sum);
        if constexpr (c.y * c.z != 0) signed_add(syz, c.y * c.z, sum);
    });
    const float rho = dr + 1.0f;
    // one-ulp reciprocal: u, Pi stay far inside the FP32 budget of the step
    const float inv = __builtin_amdgcn_rcpf(rho);
    const float ux = inv * jx.v + fx * 0.5f * inv;
    const float uy = inv * jy.v + fy * 0.5f * inv;
    const float uz = inv * jz.v + fz * 0.5f * inv;
    float pxx = sxx.v - rho * ux * ux + ux * fx;
    float pyy = syy.v - rho * uy * uy + uy * fy;
    float pzz = szz.v - rho * uz * uz + uz * fz;
    const float tr = pxx + pyy + pzz;
    pxx -= kCs2 * tr;
    pyy -= kCs2 * tr;
    pzz -= kCs2 * tr;
    m.rho = rho;
    m.drho = dr;
    m.ux = ux;
    m.uy = uy;
    m.uz = uz;
    m.pxx = pxx;
    m.pyy = pyy;
    m.pzz = pzz;
    m.pxy = sxy.v - rho * ux * uy + 0.5f * (ux * fy + uy * fx);
    m.pxz = sxz.v - rho * ux * uz + 0.5f * (ux * fz + uz * fx);
    m.pyz = syz.v - rho * uy * uz + 0.5f * (uy * fz + uz * fy);
    return m;
}

template <int Q, class P>
__device__ __forceinline__ Cell moments(const typename P::T (&f)[Q], float fx, float fy, float fz) {
    return moments_of<Q, P>([&](int i) { return f[i]; }, fx, fy, fz);
}

// Emits w_i * (even +- odd) for every direction through `emit(i, value)`.
// omco = 1 - omega.  With omco and F the reference's values this is
// feq + omco*fneq + F/2; a boundary reconstruction without force passes F=0.
// With c_u = 3 c.u, hc = c_u^2/2 - uu (uu = 3/2 u.u) and h2 = c c : Pi - t:
//     even = drho (1 + hc) + hc + 4.5 omco h2 + 1.5 (c_u c.F - u.F)
//          = E0 + c_u (rho c_u / 2 + 1.5 c.F) + 4.5 omco h2,
//       E0 = drho - rho uu - 1.5 u.F;
//     odd  = rho c_u (1 [+ c_u^2/6 - uu]) + 1.5 c.F [+ 4.5 omco h3],
// the bracketed terms for D3Q27 (D3Q27.cuh:175-184 for h3).  The per-cell
// sums are formed once, so a direction pair costs 7-10 operations.
template <int Q, class Emit>
__device__ __forceinline__ void collide(const Cell &m, float omco, float fx, float fy, float fz, Emit &&emit) {
    const float ux = m.ux, uy = m.uy, uz = m.uz;
    const float uu = 1.5f * (ux * ux + uy * uy + uz * uz);
    const float uF = ux * fx + uy * fy + uz * fz;
    const float t = kCs2 * (m.pxx + m.pyy + m.pzz);
    const float o45 = 4.5f * omco;
    const float rho = m.rho, hr = 0.5f * m.rho;
    const float E0 = __builtin_fmaf(-1.5f, uF, __builtin_fmaf(-rho, uu, m.drho));
    const float u3[3] = {3.0f * ux, 3.0f * uy, 3.0f * uz};
    const float F15[3] = {1.5f * fx, 1.5f * fy, 1.5f * fz};
    const float D[3] = {m.pxx - t, m.pyy - t, m.pzz - t};  // c_a c_a - t, one axis
    const float P2[3] = {2.0f * m.pxy, 2.0f * m.pxz, 2.0f * m.pyz};
    // third-order coefficient combinations (D3Q27.cuh:175-184)
    float axxy = 0.f, axxz = 0.f, axyy = 0.f, ayyz = 0.f, axzz = 0.f, ayzz = 0.f, axyz = 0.f;
    if constexpr (Q == 27) {
        axxy = m.pxx * uy + 2.0f * ux * m.pxy;
        axxz = m.pxx * uz + 2.0f * ux * m.pxz;
        axyy = m.pyy * ux + 2.0f * uy * m.pxy;
        ayyz = m.pyy * uz + 2.0f * uy * m.pyz;
        axzz = m.pzz * ux + 2.0f * uz * m.pxz;
        ayzz = m.pzz * uy + 2.0f * uz * m.pyz;
        axyz = ux * m.pyz + uy * m.pxz + uz * m.pxy;
    }
    // rest population: even with c = 0
    emit(0, weight<Q>(0) * __builtin_fmaf(-o45, t, E0));
    sfor<0, (Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * P + 1;
        constexpr Dir c = kDir[i];
        constexpr float w = weight<Q>(i);
        constexpr int cv[3] = {c.x, c.y, c.z};
        float cu = 0.f, cF = 0.f, h2 = 0.f;
        bool first = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (cv[a] == 0) continue;
            const float su = cv[a] > 0 ? u3[a] : -u3[a], sf = cv[a] > 0 ? F15[a] : -F15[a];
            // h2 = sum over the nonzero axes of (p_aa) - t: the first axis
            // brings D = p_aa - t, the others p_aa
            const float pa = a == 0 ? m.pxx : (a == 1 ? m.pyy : m.pzz);
            cu = first ? su : cu + su;
            cF = first ? sf : cF + sf;
            h2 = first ? D[a] : h2 + pa;
            first = false;
        }
        if constexpr (c.x * c.y != 0) h2 = __builtin_fmaf(float(c.x * c.y), P2[0], h2);
        if constexpr (c.x * c.z != 0) h2 = __builtin_fmaf(float(c.x * c.z), P2[1], h2);
        if constexpr (c.y * c.z != 0) h2 = __builtin_fmaf(float(c.y * c.z), P2[2], h2);
        const float even = __builtin_fmaf(cu, __builtin_fmaf(hr, cu, cF), __builtin_fmaf(o45, h2, E0));
        float odd;
        if constexpr (Q == 27) {
            // 3 * (c_a c_a c_b - cs2 c_b) and 6 c_x c_y c_z, all integers
            constexpr int kxxy = 3 * c.x * c.x * c.y - c.y;
            constexpr int kxxz = 3 * c.x * c.x * c.z - c.z;
            constexpr int kxyy = 3 * c.x * c.y * c.y - c.x;
            constexpr int kyyz = 3 * c.y * c.y * c.z - c.z;
            constexpr int kxzz = 3 * c.x * c.z * c.z - c.x;
            constexpr int kyzz = 3 * c.y * c.z * c.z - c.y;
            constexpr int kxyz = 6 * c.x * c.y * c.z;
            float h3 = 0.f;
            if constexpr (kxxy != 0) h3 = __builtin_fmaf(float(kxxy), axxy, h3);
            if constexpr (kxxz != 0) h3 = __builtin_fmaf(float(kxxz), axxz, h3);
            if constexpr (kxyy != 0) h3 = __builtin_fmaf(float(kxyy), axyy, h3);
            if constexpr (kyyz != 0) h3 = __builtin_fmaf(float(kyyz), ayyz, h3);
            if constexpr (kxzz != 0) h3 = __builtin_fmaf(float(kxzz), axzz, h3);
            if constexpr (kyzz != 0) h3 = __builtin_fmaf(float(kyzz), ayzz, h3);
            if constexpr (kxyz != 0) h3 = __builtin_fmaf(float(kxyz), axyz, h3);
            const float rc = rho * cu;
            odd = __builtin_fmaf(o45, h3, __builtin_fmaf(rc, __builtin_fmaf(cu * (1.0f / 6.0f), cu, -uu), rc + cF));
        } else {
            odd = __builtin_fmaf(rho, cu, cF);
        }
        const float we = w * even;
        emit(i, __builtin_fmaf(w, odd, we));
        emit(i + 1, __builtin_fmaf(-w, odd, we));
    });
}

// D3Q7 post-collision phase populations (D3Q7.cuh:124-148, lbm.cu:256-269):
//     g_k = w_k phi (1 + 4 c_k.u) + w_k gamma phi (1-phi) c_k.n = w_k (phi + c_k.o),
//     o = 4 phi u + gamma phi (1-phi) n.
// All seven are fixed by (phi, o), which is what is stored: emit(0, phi),
// emit(1 + a, o_a) for a = x, y, z.  The arriving sum rebuilds population
// +a as w (phi + o_a) and -a as w (phi - o_a), the operations the reference's
// per-population form performs on the same phi and o_a.
template <class Emit>
__device__ __forceinline__ void phase_moments(float phi, float sharp, float ux, float uy, float uz, float nx,
                                              float ny, float nz, Emit &&emit) {
    emit(0, phi);
    const float u[3] = {ux, uy, uz};
    const float n[3] = {nx, ny, nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) emit(1 + a, 4.0f * phi * u[a] + sharp * n[a]);
}

}  // namespace ssmd
