// closure.hip -- the learned subgrid closure on the gfx950 matrix cores.
//
// Network: 8 inputs -> 64 -> 64 -> 64 (ReLU) -> 1.  The three hidden layers
// run on v_mfma_f32_32x32x16_bf16 (BF16 operands, FP32 accumulation); the
// 64 -> 1 output is a per-lane dot product.  There is no counterpart in the
// reference (it has no subgrid model).
//
// Orientation: hidden units are the M rows of the matrix products and cells
// are the N columns, 32 cells per wavefront tile.  Operand layouts of the
// 32x32x16 instruction (lane l, element j of its 8-element slot, register r
// of the 16-float accumulator):
//   A[m][k]: m = l % 32, k = (l / 32) * 8 + j
//   B[k][n]: n = l % 32, k = (l / 32) * 8 + j
//   D[m][n]: n = l % 32, m = (r / 4) * 8 + (l / 32) * 4 + r % 4
// A lane therefore holds, for its cell, hidden units crow(mb, r, l) of the
// layer output.  The next layer consumes those registers directly as its B
// operand: K-chunk kc takes registers 8 (kc % 2) .. 8 (kc % 2) + 7 of block
// kc / 2, so the K index (kc, l / 32, j) names hidden unit krow(kc, l, j).
// That permutation is folded into the packed weights (A operands hold
// W[m][krow]), so no data moves between layers.
//
// Training reduces over cells, which puts cells on the K axis of the weight-
// gradient products: the deltas and activations of a tile are transposed
// through LDS for those, one wavefront per block.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/ssmd.h"
#include "lattice.hpp"

namespace ssmd {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

constexpr int IN = SSMD_MLP_INPUTS, H = SSMD_MLP_HIDDEN, NPARAM = SSMD_MLP_PARAMS;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H, OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + H;
static_assert(OFF_B4 + 1 == NPARAM, "parameter count");

// Packed network (byte offsets): one 16-byte bf16x8 slot per lane for the
// A operands, one 64-byte f32x16 accumulator image per lane for the biases
// and W4, so every lane loads its operands with vector loads.
namespace off {
constexpr size_t A1 = 0;                  // W1:  [mb 2][lane 64] bf16x8 (K 8..15 zero)
constexpr size_t A2 = A1 + 2 * 64 * 16;   // W2:  [mb 2][kc 4][lane 64] bf16x8
constexpr size_t A3 = A2 + 8 * 64 * 16;   // W3
constexpr size_t T3 = A3 + 8 * 64 * 16;   // W3^T (backward)
constexpr size_t T2 = T3 + 8 * 64 * 16;   // W2^T (backward)
constexpr size_t B1 = T2 + 8 * 64 * 16;   // b1:  [mb 2][lane 64] f32x16
constexpr size_t B2 = B1 + 2 * 64 * 64;
constexpr size_t B3 = B2 + 2 * 64 * 64;
constexpr size_t W4 = B3 + 2 * 64 * 64;   // W4 in the accumulator image
constexpr size_t B4 = W4 + 2 * 64 * 64;   // b4, 3 floats of padding
constexpr size_t BYTES = B4 + 16;
}  // namespace off

// Hidden unit held in accumulator register r of block mb by lane l.
__host__ __device__ constexpr int crow(int mb, int r, int l) { return mb * 32 + (r / 4) * 8 + (l / 32) * 4 + r % 4; }
// Hidden unit at K index (kc, l / 32, j) of a B operand built from accumulators.
__host__ __device__ constexpr int krow(int kc, int l, int j) { return crow(kc / 2, 8 * (kc % 2) + j, l); }

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Two floats -> two bf16 (round to nearest even) in one word, low first:
// one two-operand v_cvt_pk_bf16_f32 (two scalar conversions would take two
// of them and a v_perm_b32).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
// max(x, 0) of a finite float on its bit pattern: negative floats (sign bit
// set, -0 included) are negative integers.  One v_max_i32; fmaxf would add a
// canonicalizing v_max_f32.
__device__ __forceinline__ float relu(float x) {
    const int v = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(float, v > 0 ? v : 0);
}
// Same, then ReLU on the pair as 16-bit integers: a bf16 is negative exactly
// when its sign bit is set, and +0 is the zero pattern.
__device__ __forceinline__ uint32_t pk_relu(float a, float b) {
    s16x2 v = __builtin_bit_cast(s16x2, pk(a, b));
    v = __builtin_elementwise_max(v, s16x2{0, 0});
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf16_round(float v) {
    const uint32_t u = uint32_t(__builtin_bit_cast(uint16_t, static_cast<__bf16>(v))) << 16;
    return __builtin_bit_cast(float, u);
}

// B operand of K-chunk KC from the accumulators of the previous layer.
template <int KC, bool RELU>
__device__ __forceinline__ bf16x8 chunk(const f32x16 (&c)[2]) {
    constexpr int MB = KC / 2, R0 = 8 * (KC % 2);
    u32x4 w;
    if constexpr (RELU) {
        w.x = pk_relu(c[MB][R0 + 0], c[MB][R0 + 1]);
        w.y = pk_relu(c[MB][R0 + 2], c[MB][R0 + 3]);
        w.z = pk_relu(c[MB][R0 + 4], c[MB][R0 + 5]);
        w.w = pk_relu(c[MB][R0 + 6], c[MB][R0 + 7]);
    } else {
        w.x = pk(c[MB][R0 + 0], c[MB][R0 + 1]);
        w.y = pk(c[MB][R0 + 2], c[MB][R0 + 3]);
        w.z = pk(c[MB][R0 + 4], c[MB][R0 + 5]);
        w.w = pk(c[MB][R0 + 6], c[MB][R0 + 7]);
    }
    return __builtin_bit_cast(bf16x8, w);
}

// c = W * b + bias for a 64-unit layer with K = 64 (four chunks of B).
template <class AOp>
__device__ __forceinline__ void layer64(AOp &&a, const bf16x8 (&b)[4], const f32x16 (&bias)[2], f32x16 (&c)[2]) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        f32x16 acc = bias[mb];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) acc = mfma(a(mb, kc), b[kc], acc);
        c[mb] = acc;
    }
}

template <bool RELU>
__device__ __forceinline__ void chunks(const f32x16 (&c)[2], bf16x8 (&b)[4]) {
    b[0] = chunk<0, RELU>(c);
    b[1] = chunk<1, RELU>(c);
    b[2] = chunk<2, RELU>(c);
    b[3] = chunk<3, RELU>(c);
}

// Network output of the tile (lanes l and l ^ 32 both get cell l % 32).
__device__ __forceinline__ float out_dot(const f32x16 (&c3)[2], const f32x16 (&w4)[2], float b4) {
    float s = 0.f;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) s += relu(c3[mb][r]) * w4[mb][r];
    return s + __shfl_xor(s, 32) + b4;
}

__device__ __forceinline__ bf16x8 load_x(const uint16_t *__restrict__ x, uint32_t cell, bool ok) {
    bf16x8 v = {};
    if (ok) v = *reinterpret_cast<const bf16x8 *>(x + size_t(cell) * IN);
    return v;
}

// --------------------------------------------------------------------------
// Inference: each wavefront keeps the weight operands in registers (72
// VGPRs) and walks 32-cell tiles.  The biases and W4 live once per block in
// LDS in unit order (1 KB): a lane's accumulator image of one block mb is
// four runs of four consecutive units, mb*32 + j*8 + (l/32)*4 + 0..3, one
// broadcast 16-byte LDS read each, so they initialize the accumulators
// without holding 128 VGPRs of replicated values (3 waves per SIMD instead
// of 1).  closure != 0 writes max(floor, scale * y) (the eddy relaxation
// time), else y.
__device__ __forceinline__ f32x16 unit_image(const float *v, int mb, int lane) {
    const int base = mb * 32 + (lane / 32) * 4;
    f32x16 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float4 q = *reinterpret_cast<const float4 *>(v + base + j * 8);
        r[4 * j + 0] = q.x;
        r[4 * j + 1] = q.y;
        r[4 * j + 2] = q.z;
        r[4 * j + 3] = q.w;
    }
    return r;
}

// (2 waves per SIMD: the two-tile loop holds the weights, two tiles'
// accumulators and their B operands in 200 VGPRs; left free, the compiler
// parks accumulators in AGPRs and runs 1 wave per SIMD)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_forward(const uint8_t *__restrict__ packed, const uint16_t *__restrict__ x,
                                                 float *__restrict__ y, uint32_t n, float scale, float floor,
                                                 int closure) {
    // b1, b2, b3, W4 in unit order, read back from the accumulator images:
    // lane 0 holds units mb*32 + j*8 + i, lane 32 units mb*32 + j*8 + 4 + i
    // (i = 0..3) in register 4j + i
    __shared__ __attribute__((aligned(16))) float units[4][H];
    {
        const float *img = reinterpret_cast<const float *>(packed + off::B1);
        const int t = int(threadIdx.x), a = t / H, u = t % H;  // 256 threads = 4 x 64 entries
        const int mb = u / 32, rem = u % 32, j = rem / 8, half = (rem % 8) / 4, i = rem % 4;
        units[a][u] = img[(size_t(a * 2 + mb) * 64 + half * 32) * 16 + 4 * j + i];
    }
    __syncthreads();
    const int lane = int(threadIdx.x) & 63;
    const bf16x8 *A1 = reinterpret_cast<const bf16x8 *>(packed + off::A1);
    const bf16x8 *A2 = reinterpret_cast<const bf16x8 *>(packed + off::A2);
    const bf16x8 *A3 = reinterpret_cast<const bf16x8 *>(packed + off::A3);
    bf16x8 a1[2], a2[2][4], a3[2][4];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        a1[mb] = A1[mb * 64 + lane];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            a2[mb][kc] = A2[(mb * 4 + kc) * 64 + lane];
            a3[mb][kc] = A3[(mb * 4 + kc) * 64 + lane];
        }
    }
    const float b4 = *reinterpret_cast<const float *>(packed + off::B4);

    const uint32_t tiles = (n + 31) / 32;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = gridDim.x * blockDim.x / 64;
    // Two tiles per iteration, A and B, in an order that puts one tile's VALU
    // epilogue (bf16 packing, ReLU, the output dot product) next to the other
    // tile's matrix products, which the two pipes can then overlap.  Each
    // tile's arithmetic is that of a single tile.
    for (uint32_t t = wave; t < tiles; t += 2 * nwaves) {
        const uint32_t cellA = t * 32 + uint32_t(lane & 31), cellB = cellA + nwaves * 32;
        const bool okA = cellA < n, okB = cellB < n;
        const bf16x8 xA = load_x(x, cellA, lane < 32 && okA), xB = load_x(x, cellB, lane < 32 && okB);
        // an offset the compiler cannot see is 0: the LDS reads below stay in
        // the loop instead of being hoisted into 128 registers
        int z0 = 0;
        asm volatile("" : "+v"(z0));
        const float *u0 = &units[0][0] + z0;
        f32x16 cA[2], cB[2];
        bf16x8 bA[4], bB[4];
        auto hidden = [&](const bf16x8 (&a)[2][4], const float *bias, const bf16x8 (&b)[4], f32x16 (&c)[2]) {
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                f32x16 acc = unit_image(bias, mb, lane);
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) acc = mfma(a[mb][kc], b[kc], acc);
                c[mb] = acc;
            }
        };
        // output: relu(c) . W4, four units of W4 at a time from LDS
        auto output = [&](const f32x16 (&c)[2]) {
            float o = 0.f;
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float4 w = *reinterpret_cast<const float4 *>(u0 + 3 * H + mb * 32 + (lane / 32) * 4 + j * 8);
                    o += relu(c[mb][4 * j + 0]) * w.x;
                    o += relu(c[mb][4 * j + 1]) * w.y;
                    o += relu(c[mb][4 * j + 2]) * w.z;
                    o += relu(c[mb][4 * j + 3]) * w.w;
                }
            return o + __shfl_xor(o, 32) + b4;  // the association of out_dot
        };
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) cA[mb] = mfma(a1[mb], xA, unit_image(u0, mb, lane));
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) cB[mb] = mfma(a1[mb], xB, unit_image(u0, mb, lane));
        chunks<true>(cA, bA);
        hidden(a2, u0 + H, bA, cA);
        chunks<true>(cB, bB);
        hidden(a2, u0 + H, bB, cB);
        chunks<true>(cA, bA);
        hidden(a3, u0 + 2 * H, bA, cA);
        chunks<true>(cB, bB);
        hidden(a3, u0 + 2 * H, bB, cB);
        const float oA = output(cA);
        const float oB = output(cB);
        if (lane < 32 && okA) y[cellA] = closure ? __builtin_fmaxf(scale * oA, floor) : oA;
        if (lane < 32 && okB) y[cellB] = closure ? __builtin_fmaxf(scale * oB, floor) : oB;
    }
}

// Blocks of k_forward for n cells: at most the blocks resident at once, so
// each wavefront loads the weight operands once and walks many tiles.
inline uint32_t forward_blocks(uint32_t n) {
    static uint32_t resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_forward, 256, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        resident = uint32_t(cus) * uint32_t(per_cu);
    }
    const uint32_t tiles = (n + 31) / 32, want = (tiles + 3) / 4;
    return want < resident ? want : resident;
}

// --------------------------------------------------------------------------
// Training: forward and backward of the mean squared error for 32-sample
// tiles, one wavefront per block, gradient partials per block into `work`
// (row of NPARAM + 1 floats: the gradient, then the weighted sum of squared
// errors).  Sample weights wt may be null (all 1).
struct Tr {
    float act[H][33];  // right-hand activations of a weight gradient, [unit][cell]
    float dz[H][33];   // bf16-rounded deltas, [unit][cell]
    float dy[32];
};

__device__ __forceinline__ void store_t(float (&buf)[H][33], const f32x16 (&c)[2], int lane, bool relu_round) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float v = c[mb][r];
            if (relu_round) v = bf16_round(__builtin_fmaxf(v, 0.f));
            buf[crow(mb, r, lane)][lane & 31] = v;
        }
}
// bf16x8 of row (blk * 32 + l % 32), cells kc * 16 + (l / 32) * 8 .. + 7
__device__ __forceinline__ bf16x8 rowchunk(const float (&buf)[H][33], int blk, int kc, int lane) {
    const float *p = &buf[blk * 32 + (lane & 31)][kc * 16 + (lane / 32) * 8];
    u32x4 w;
    w.x = pk(p[0], p[1]);
    w.y = pk(p[2], p[3]);
    w.z = pk(p[4], p[5]);
    w.w = pk(p[6], p[7]);
    return __builtin_bit_cast(bf16x8, w);
}
__device__ __forceinline__ float rowsum(const float (&buf)[H][33], int lane) {
    float s = 0.f;
#pragma unroll 8
    for (int c = 0; c < 32; ++c) s += buf[lane][c];
    return s;
}

__global__ void __launch_bounds__(64) k_grad(const uint8_t *__restrict__ packed, const uint16_t *__restrict__ x,
                                            const float *__restrict__ t, const float *__restrict__ wt, uint32_t n,
                                            float *__restrict__ work) {
    __shared__ Tr sm;
    const int lane = int(threadIdx.x);
    const bf16x8 *A1 = reinterpret_cast<const bf16x8 *>(packed + off::A1);
    const bf16x8 *A2 = reinterpret_cast<const bf16x8 *>(packed + off::A2);
    const bf16x8 *A3 = reinterpret_cast<const bf16x8 *>(packed + off::A3);
    const bf16x8 *T3 = reinterpret_cast<const bf16x8 *>(packed + off::T3);
    const bf16x8 *T2 = reinterpret_cast<const bf16x8 *>(packed + off::T2);
    const float b4 = *reinterpret_cast<const float *>(packed + off::B4);
    const float two_over_n = 2.0f / float(n);

    f32x16 g1[2] = {}, g2[2][2] = {}, g3[2][2] = {};
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f, gb4 = 0.f, sse = 0.f;
    const bf16x8 zero8 = {};

    const uint32_t tiles = (n + 31) / 32;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t cell = tile * 32 + uint32_t(lane & 31);
        const bool ok = cell < n;
        const bf16x8 xin = load_x(x, cell, lane < 32 && ok);
        // biases and W4 are reloaded per tile (L1/L2 hits): the base is made
        // opaque so they are not held in registers across the loop
        const uint8_t *pk_base = packed;
        asm volatile("" : "+s"(pk_base));
        const f32x16 *Bs = reinterpret_cast<const f32x16 *>(pk_base + off::B1);
        f32x16 b1[2], b2[2], b3[2], w4[2];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            b1[mb] = Bs[(0 * 2 + mb) * 64 + lane];
            b2[mb] = Bs[(1 * 2 + mb) * 64 + lane];
            b3[mb] = Bs[(2 * 2 + mb) * 64 + lane];
            w4[mb] = Bs[(3 * 2 + mb) * 64 + lane];
        }
        // forward, keeping the pre-activations
        f32x16 c1[2], c2[2], c3[2];
        bf16x8 b[4];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) c1[mb] = mfma(A1[mb * 64 + lane], xin, b1[mb]);
        chunks<true>(c1, b);
        layer64([&](int mb, int kc) { return A2[(mb * 4 + kc) * 64 + lane]; }, b, b2, c2);
        chunks<true>(c2, b);
        layer64([&](int mb, int kc) { return A3[(mb * 4 + kc) * 64 + lane]; }, b, b3, c3);
        const float yv = out_dot(c3, w4, b4);
        const float e = ok ? yv - t[cell] : 0.f;
        const float we = ok && wt ? wt[cell] * e : e;
        const float dy = we * two_over_n;
        if (lane < 32) sse += we * e;

        // output layer: dW4 = sum dy h3, db4 = sum dy; dz3 = [z3 > 0] dy W4
        f32x16 d[2];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float z = c3[mb][r];
                sm.dz[crow(mb, r, lane)][lane & 31] = __builtin_fmaxf(z, 0.f) * dy;
                d[mb][r] = z > 0.f ? dy * w4[mb][r] : 0.f;
            }
        if (lane < 32) sm.dy[lane] = dy;
        __syncthreads();
        gw4 += rowsum(sm.dz, lane);
        if (lane == 0)
            for (int c = 0; c < 32; ++c) gb4 += sm.dy[c];
        __syncthreads();

        // layer 3: dW3 = q(dz3) h2^T, db3 = sum q(dz3); dz2 = [z2 > 0] W3^T q(dz3)
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) d[mb][r] = bf16_round(d[mb][r]);
        store_t(sm.dz, d, lane, false);
        store_t(sm.act, c2, lane, true);
        __syncthreads();
        gb3 += rowsum(sm.dz, lane);
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int kc = 0; kc < 2; ++kc)
                    g3[mb][nb] = mfma(rowchunk(sm.dz, mb, kc, lane), rowchunk(sm.act, nb, kc, lane), g3[mb][nb]);
        {
            chunks<false>(d, b);
            const f32x16 z16[2] = {};
            layer64([&](int mb, int kc) { return T3[(mb * 4 + kc) * 64 + lane]; }, b, z16, d);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int r = 0; r < 16; ++r) d[mb][r] = c2[mb][r] > 0.f ? bf16_round(d[mb][r]) : 0.f;
        }
        __syncthreads();

        // layer 2: dW2 = q(dz2) h1^T, db2; dz1 = [z1 > 0] W2^T q(dz2)
        store_t(sm.dz, d, lane, false);
        store_t(sm.act, c1, lane, true);
        __syncthreads();
        gb2 += rowsum(sm.dz, lane);
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int kc = 0; kc < 2; ++kc)
                    g2[mb][nb] = mfma(rowchunk(sm.dz, mb, kc, lane), rowchunk(sm.act, nb, kc, lane), g2[mb][nb]);
        {
            chunks<false>(d, b);
            const f32x16 z16[2] = {};
            layer64([&](int mb, int kc) { return T2[(mb * 4 + kc) * 64 + lane]; }, b, z16, d);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int r = 0; r < 16; ++r) d[mb][r] = c1[mb][r] > 0.f ? bf16_round(d[mb][r]) : 0.f;
        }
        __syncthreads();

        // layer 1: dW1 = q(dz1) x^T, db1
        store_t(sm.dz, d, lane, false);
        if (lane < 32) {
            const u32x4 w = __builtin_bit_cast(u32x4, xin);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sm.act[2 * j][lane] = __builtin_bit_cast(float, ws[j] << 16);
                sm.act[2 * j + 1][lane] = __builtin_bit_cast(float, ws[j] & 0xffff0000u);
            }
        }
        __syncthreads();
        gb1 += rowsum(sm.dz, lane);
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int kc = 0; kc < 2; ++kc)
                g1[mb] = mfma(rowchunk(sm.dz, mb, kc, lane), (lane & 31) < IN ? rowchunk(sm.act, 0, kc, lane) : zero8,
                              g1[mb]);
        __syncthreads();
    }

    float *o = work + size_t(blockIdx.x) * (NPARAM + 1);
    const int nn = lane & 31;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = mb * 32 + (r / 4) * 8 + (lane / 32) * 4 + r % 4;
            if (nn < IN) o[OFF_W1 + m * IN + nn] = g1[mb][r];
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                o[OFF_W2 + m * H + nb * 32 + nn] = g2[mb][nb][r];
                o[OFF_W3 + m * H + nb * 32 + nn] = g3[mb][nb][r];
            }
        }
    o[OFF_B1 + lane] = gb1;
    o[OFF_B2 + lane] = gb2;
    o[OFF_B3 + lane] = gb3;
    o[OFF_W4 + lane] = gw4;
    for (int s = 32; s >= 1; s >>= 1) sse += __shfl_xor(sse, s);
    if (lane == 0) {
        o[OFF_B4] = gb4;
        o[NPARAM] = sse;
    }
}

// Sum of the per-block partials in block order (deterministic), in double.
__global__ void __launch_bounds__(256) k_grad_reduce(const float *__restrict__ work, uint32_t blocks, uint32_t n,
                                                     float *__restrict__ grad, float *__restrict__ loss) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i > uint32_t(NPARAM)) return;
    double s = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) s += double(work[size_t(b) * (NPARAM + 1) + i]);
    if (i < uint32_t(NPARAM))
        grad[i] = float(s);
    else
        *loss = float(s / double(n));
}

__global__ void __launch_bounds__(256) k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, float lr, float b1, float b2, float eps, float c1,
                                              float c2) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= uint32_t(NPARAM)) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + eps);
}

// params -> packed operand slots; one thread per packed element.
constexpr uint32_t kPackA = 2 * 64 * 8 + 4 * (8 * 64 * 8);  // bf16 elements: A1, A2, A3, T3, T2
constexpr uint32_t kPackF = 4 * (2 * 64 * 16) + 4;          // floats: B1, B2, B3, W4, B4(+pad)

__global__ void __launch_bounds__(256) k_pack(const float *__restrict__ p, uint8_t *__restrict__ packed) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < kPackA) {
        float v = 0.f;
        size_t byte;
        if (i < 2 * 64 * 8) {  // A1: W1[mb*32 + l%32][j] in the lower half, 0 in the upper
            const int j = int(i % 8), l = int(i / 8 % 64), mb = int(i / 512);
            if (l < 32) v = p[OFF_W1 + (mb * 32 + l) * IN + j];
            byte = off::A1 + size_t(i) * 2;
        } else {
            const uint32_t r = i - 2 * 64 * 8;
            const int sec = int(r / 4096), e = int(r % 4096);
            const int j = e % 8, l = e / 8 % 64, kc = e / 512 % 4, mb = e / 2048;
            const int m = mb * 32 + l % 32, k = krow(kc, l, j);
            switch (sec) {
                case 0: v = p[OFF_W2 + m * H + k]; byte = off::A2; break;
                case 1: v = p[OFF_W3 + m * H + k]; byte = off::A3; break;
                case 2: v = p[OFF_W3 + k * H + m]; byte = off::T3; break;
                default: v = p[OFF_W2 + k * H + m]; byte = off::T2; break;
            }
            byte += size_t(e) * 2;
        }
        *reinterpret_cast<uint16_t *>(packed + byte) = __builtin_bit_cast(uint16_t, static_cast<__bf16>(v));
        return;
    }
    const uint32_t f = i - kPackA;
    if (f >= kPackF) return;
    float v = 0.f;
    if (f < 4 * 2 * 64 * 16) {
        const int sec = int(f / 2048), e = int(f % 2048);
        const int r = e % 16, l = e / 16 % 64, mb = e / 1024;
        const int u = crow(mb, r, l);
        const int base = sec == 0 ? OFF_B1 : (sec == 1 ? OFF_B2 : (sec == 2 ? OFF_B3 : OFF_W4));
        v = p[base + u];
    } else if (f == 4 * 2 * 64 * 16) {
        v = p[OFF_B4];
    }
    reinterpret_cast<float *>(packed + off::B1)[f] = v;
}

}  // namespace mlp

// --------------------------------------------------------------------------
// Training data: a fine field set box-filtered by 2 onto the coarse grid.
// Coarse interior cell i (per axis, 1 <= i <= n_c - 2) covers the fine
// cells 2i - 1 and 2i.  With acoustic scaling (dx and dt both halve on the
// fine grid) lattice velocities agree and a strain rate per coarse step is
// twice the fine one.
namespace sgs {

enum Plane { kRho, kUx, kUy, kUz, kPhi, kS = 5, kT = 11, kJ = 17, kPlanes = 20 };  // S, T: xx yy zz xy xz yz; J: x y z

struct Fine {
    int nx, ny, nz;
    const float *rho, *ux, *uy, *uz, *pxx, *pyy, *pzz, *pxy, *pxz, *pyz, *phi;
    float visc_water, visc_diff;
    // fine -> coarse lattice units: rt = dt_c / dt_f (time), rv = rt / 2 (velocity)
    float rt, rv;
};

__device__ __forceinline__ bool inner(int x, int y, int z, int nx, int ny, int nz) {
    return x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && z >= 1 && z <= nz - 2;
}

// Per coarse cell: Favre-filtered velocity, filtered rho, phi and strain rate
// (fine S = -3 Pi / (2 rho tau), tau = 1/2 + 3 nu(phi)), and the subgrid
// stress T = <rho u u> - <rho> u~ u~ and the subgrid flux of the phase field
// J = <phi u> - <phi> <u> (plain box averages), all in coarse lattice units
// (velocity and flux x rv, strain rate x rt, stress x rv^2).  Ghost cells:
// rho 1, the rest 0.
__global__ void __launch_bounds__(256) k_filter(Fine f, int cnx, int cny, int cnz, float *__restrict__ out) {
    const size_t nc = size_t(cnx) * cny * cnz;
    const size_t c = size_t(blockIdx.x) * 256u + threadIdx.x;
    if (c >= nc) return;
    const int cx = int(c % cnx), cy = int(c / cnx % cny), cz = int(c / (size_t(cnx) * cny));
    float v[kPlanes];
#pragma unroll
    for (int i = 0; i < kPlanes; ++i) v[i] = 0.f;
    if (!inner(cx, cy, cz, cnx, cny, cnz)) {
        v[kRho] = 1.f;
    } else {
        float r = 0, mu[3] = {0, 0, 0}, muu[6] = {0, 0, 0, 0, 0, 0}, ph = 0, st[6] = {0, 0, 0, 0, 0, 0};
        float us[3] = {0, 0, 0}, pu[3] = {0, 0, 0};  // sum u, sum phi u
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const size_t q = size_t(2 * cx - 1 + dx) +
                                     size_t(f.nx) * (size_t(2 * cy - 1 + dy) + size_t(f.ny) * size_t(2 * cz - 1 + dz));
                    const float rho = f.rho[q], u[3] = {f.ux[q], f.uy[q], f.uz[q]}, p = f.phi[q];
                    const float tau = 0.5f + 3.0f * __builtin_fmaf(p, f.visc_diff, f.visc_water);
                    const float sc = -1.5f / (rho * tau);
                    const float pi[6] = {f.pxx[q], f.pyy[q], f.pzz[q], f.pxy[q], f.pxz[q], f.pyz[q]};
                    r += rho;
                    ph += p;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        mu[a] += rho * u[a];
                        us[a] += u[a];
                        pu[a] += p * u[a];
                    }
                    muu[0] += rho * u[0] * u[0];
                    muu[1] += rho * u[1] * u[1];
                    muu[2] += rho * u[2] * u[2];
                    muu[3] += rho * u[0] * u[1];
                    muu[4] += rho * u[0] * u[2];
                    muu[5] += rho * u[1] * u[2];
#pragma unroll
                    for (int a = 0; a < 6; ++a) st[a] += sc * pi[a];
                }
        const float rb = r * 0.125f, ut[3] = {mu[0] / r, mu[1] / r, mu[2] / r};
        v[kRho] = rb;
        v[kUx] = f.rv * ut[0];
        v[kUy] = f.rv * ut[1];
        v[kUz] = f.rv * ut[2];
        v[kPhi] = ph * 0.125f;
#pragma unroll
        for (int a = 0; a < 3; ++a) v[kJ + a] = f.rv * (pu[a] * 0.125f - (ph * 0.125f) * (us[a] * 0.125f));
        const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
        const float rv2 = f.rv * f.rv;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            v[kS + a] = f.rt * st[a] * 0.125f;
            v[kT + a] = rv2 * (muu[a] * 0.125f - rb * ut[ia[a]] * ut[ib[a]]);
        }
    }
#pragma unroll
    for (int i = 0; i < kPlanes; ++i) out[size_t(i) * nc + c] = v[i];
}

// Network inputs (as the LEARNED step writes them), the least-squares eddy
// relaxation time of the subgrid stress and its weight, per coarse cell:
//   nu_t = -(dev T : S) / (2 rho S:S),  target = 3 nu_t / tau_scale,
//   weight = S:S (so the weighted squared error is the stress residual's).
// Ghost cells get weight 0.
template <int Q>
__global__ void __launch_bounds__(256) k_samples(int cnx, int cny, int cnz, const float *__restrict__ in,
                                                 float strain_scale, float grad_scale, float inv_tau_scale,
                                                 uint16_t *__restrict__ feat, float *__restrict__ target,
                                                 float *__restrict__ weight) {
    const size_t nc = size_t(cnx) * cny * cnz;
    const size_t c = size_t(blockIdx.x) * 256u + threadIdx.x;
    if (c >= nc) return;
    const int cx = int(c % cnx), cy = int(c / cnx % cny), cz = int(c / (size_t(cnx) * cny));
    uint4 w = {0u, 0u, 0u, 0u};
    float t = 0.f, wt = 0.f;
    if (inner(cx, cy, cz, cnx, cny, cnz)) {
        const float *phi = in + size_t(kPhi) * nc;
        auto at = [&](int dx, int dy, int dz) {
            const int x = cx + dx, y = cy + dy, z = cz + dz;
            return inner(x, y, z, cnx, cny, cnz) ? phi[size_t(x) + size_t(cnx) * (size_t(y) + size_t(cny) * z)]
                                                 : 0.f;
        };
        float gx, gy, gz;
        iso_grad<Q>(at, gx, gy, gz);
        const float g = __builtin_sqrtf(gx * gx + gy * gy + gz * gz);
        float S[6], T[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            S[a] = in[size_t(kS + a) * nc + c];
            T[a] = in[size_t(kT + a) * nc + c];
        }
        const float tr = (T[0] + T[1] + T[2]) * (1.0f / 3.0f);
        const float ss = S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2.0f * (S[3] * S[3] + S[4] * S[4] + S[5] * S[5]);
        const float ts = (T[0] - tr) * S[0] + (T[1] - tr) * S[1] + (T[2] - tr) * S[2] +
                         2.0f * (T[3] * S[3] + T[4] * S[4] + T[5] * S[5]);
        const float rho = in[size_t(kRho) * nc + c];
        if (ss > 0.f) {
            t = 3.0f * (-ts / (2.0f * rho * ss)) * inv_tau_scale;
            wt = ss;
        }
        w.x = mlp::pk(S[0] * strain_scale, S[1] * strain_scale);
        w.y = mlp::pk(S[2] * strain_scale, S[3] * strain_scale);
        w.z = mlp::pk(S[4] * strain_scale, S[5] * strain_scale);
        w.w = mlp::pk(in[size_t(kPhi) * nc + c], g * grad_scale);
    }
    *reinterpret_cast<uint4 *>(feat + c * 8) = w;
    target[c] = t;
    weight[c] = wt;
}

}  // namespace sgs
}  // namespace ssmd

// ---------------------------------------------------------------------------
namespace {
int herr(hipError_t e) { return e == hipSuccess ? 0 : int(e); }

// Blocks of the gradient kernel for n samples: enough 32-sample tiles per
// block to amortize the per-block partial row, at most one block per tile.
constexpr uint32_t grad_blocks(uint32_t n) {
    const uint32_t tiles = (n + 31) / 32;
    uint32_t b = (tiles + 7) / 8;
    if (b > 2048) b = 2048;
    return b < 1 ? 1 : b;
}

// Workspace of ssmd_mlp_grad_noalloc: the per-block gradient partials for up
// to SSMD_MLP_GRAD_NOALLOC_MAX samples, one copy per device for the life of
// the process.  k_grad writes every element of the rows it uses, so nothing
// needs clearing between calls.
constexpr uint32_t kNoallocBlocks = grad_blocks(SSMD_MLP_GRAD_NOALLOC_MAX);
constexpr size_t kNoallocFloats = size_t(kNoallocBlocks) * (SSMD_MLP_PARAMS + 1);
static __device__ __align__(256) float scratch[kNoallocFloats];

// Device address of `scratch` per device (the host-side name is only a
// stand-in for it), looked up once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(float **out) {
    int dev = 0;
    if (const hipError_t e = hipGetDevice(&dev); e != hipSuccess) return int(e);
    if (dev < 0 || dev >= kMaxDevices) return SSMD_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)); e != hipSuccess) return int(e);
        if (!ptr) return SSMD_ENOTREADY;
        g_scratch_ptr[dev] = static_cast<float *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return SSMD_OK;
}
}  // namespace

extern "C" {

size_t ssmd_mlp_packed_bytes(void) { return ssmd::mlp::off::BYTES; }

int ssmd_mlp_pack(const float *params, void *packed, void *stream) {
    using namespace ssmd::mlp;
    if (!params || !packed) return SSMD_EINVAL;
    const uint32_t total = kPackA + kPackF;
    hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), params,
                       static_cast<uint8_t *>(packed));
    return herr(hipGetLastError());
}

int ssmd_mlp_forward(const void *packed, const uint16_t *x, float *y, uint32_t n, void *stream) {
    using namespace ssmd::mlp;
    if (!packed || (n && (!x || !y))) return SSMD_EINVAL;
    if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) return SSMD_EINVAL;
    if (n == 0) return SSMD_OK;
    hipLaunchKernelGGL(k_forward, dim3(forward_blocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint8_t *>(packed), x, y, n, 1.0f, 0.0f, 0);
    return herr(hipGetLastError());
}

size_t ssmd_mlp_grad_work_bytes(uint32_t n) {
    return size_t(grad_blocks(n)) * (SSMD_MLP_PARAMS + 1) * sizeof(float);
}

int ssmd_mlp_grad(const void *packed, const uint16_t *x, const float *t, const float *w, uint32_t n, float *grad,
                  float *loss, void *work, size_t work_bytes, void *stream) {
    using namespace ssmd::mlp;
    if (!packed || !x || !t || !grad || !loss || !work || n == 0) return SSMD_EINVAL;
    if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) return SSMD_EINVAL;
    if (work_bytes < ssmd_mlp_grad_work_bytes(n)) return SSMD_EINVAL;
    const uint32_t blocks = grad_blocks(n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_grad, dim3(blocks), dim3(64), 0, s, static_cast<const uint8_t *>(packed), x, t, w, n,
                       static_cast<float *>(work));
    if (const int rc = herr(hipGetLastError())) return rc;
    hipLaunchKernelGGL(k_grad_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s,
                       static_cast<const float *>(work), blocks, n, grad, loss);
    return herr(hipGetLastError());
}

int ssmd_mlp_grad_noalloc(const void *packed, const uint16_t *x, const float *t, const float *w, uint32_t n,
                          float *grad, float *loss, void *stream) {
    using namespace ssmd::mlp;
    if (!packed || !x || !t || !grad || !loss || n == 0) return SSMD_EINVAL;
    if (n > SSMD_MLP_GRAD_NOALLOC_MAX) return SSMD_EINVAL;
    if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) return SSMD_EINVAL;
    float *work = nullptr;
    if (const int rc = scratch_address(&work); rc != 0) return rc;
    const uint32_t blocks = grad_blocks(n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_grad, dim3(blocks), dim3(64), 0, s, static_cast<const uint8_t *>(packed), x, t, w, n, work);
    if (const int rc = herr(hipGetLastError())) return rc;
    hipLaunchKernelGGL(k_grad_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, work, blocks, n, grad, loss);
    return herr(hipGetLastError());
}

int ssmd_mlp_adam(float *params, const float *grad, float *m, float *v, uint32_t step, float lr, float beta1,
                  float beta2, float eps, void *stream) {
    using namespace ssmd::mlp;
    if (!params || !grad || !m || !v || step == 0) return SSMD_EINVAL;
    const float c1 = float(1.0 - __builtin_pow(double(beta1), double(step)));
    const float c2 = float(1.0 - __builtin_pow(double(beta2), double(step)));
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), params,
                       grad, m, v, lr, beta1, beta2, eps, c1, c2);
    return herr(hipGetLastError());
}

size_t ssmd_sgs_work_bytes(const ssmd_params *coarse) {
    if (!coarse) return 0;
    return size_t(ssmd::sgs::kPlanes) * coarse->nx * coarse->ny * coarse->nz * sizeof(float);
}

int ssmd_sgs_samples(const ssmd_params *fine, const ssmd_fields *fd, const ssmd_params *coarse, float *work,
                     uint16_t *feat, float *target, float *weight, void *stream) {
    using namespace ssmd::sgs;
    if (!fine || !fd || !coarse || !work || !feat || !target || !weight) return SSMD_EINVAL;
    if (ssmd_validate(fine) != 0 || ssmd_validate(coarse) != 0) return SSMD_EINVAL;
    if (coarse->q != 19 && coarse->q != 27) return SSMD_EINVAL;
    if (2 * (coarse->nx - 2) > fine->nx - 2 || 2 * (coarse->ny - 2) > fine->ny - 2 ||
        2 * (coarse->nz - 2) > fine->nz - 2)
        return SSMD_EINVAL;
    if ((reinterpret_cast<uintptr_t>(feat) & 15) != 0) return SSMD_EINVAL;
    if (!fd->rho || !fd->ux || !fd->uy || !fd->uz || !fd->pxx || !fd->pyy || !fd->pzz || !fd->pxy || !fd->pxz ||
        !fd->pyz || !fd->phi)
        return SSMD_ENOTREADY;
    Fine f{};
    f.nx = int(fine->nx);
    f.ny = int(fine->ny);
    f.nz = int(fine->nz);
    f.rho = fd->rho;
    f.ux = fd->ux;
    f.uy = fd->uy;
    f.uz = fd->uz;
    f.pxx = fd->pxx;
    f.pyy = fd->pyy;
    f.pzz = fd->pzz;
    f.pxy = fd->pxy;
    f.pxz = fd->pxz;
    f.pyz = fd->pyz;
    f.phi = fd->phi;
    f.visc_water = float(fine->visc_water);
    f.visc_diff = float(fine->visc_oil) - float(fine->visc_water);
    // The two runs describe one physical flow with dx_c = 2 dx_f, so equal
    // physical viscosity fixes the time ratio: nu_c = nu_f rt / 4.
    if (!(fine->visc_water > 0.0) || !(coarse->visc_water > 0.0)) return SSMD_EINVAL;
    const double rt = 4.0 * coarse->visc_water / fine->visc_water;
    if (!(rt >= 1.0 && rt <= 16.0)) return SSMD_EINVAL;
    f.rt = float(rt);
    f.rv = float(0.5 * rt);
    const size_t nc = size_t(coarse->nx) * coarse->ny * coarse->nz;
    const dim3 grid(unsigned((nc + 255) / 256)), block(256);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_filter, grid, block, 0, s, f, int(coarse->nx), int(coarse->ny), int(coarse->nz), work);
    if (const int rc = herr(hipGetLastError())) return rc;
    const float ss = float(coarse->strain_scale), gs = float(coarse->grad_scale);
    const float its = coarse->tau_scale > 0.0 ? float(1.0 / coarse->tau_scale) : 1.0f;
    if (coarse->q == 19)
        hipLaunchKernelGGL(k_samples<19>, grid, block, 0, s, int(coarse->nx), int(coarse->ny), int(coarse->nz), work,
                           ss, gs, its, feat, target, weight);
    else
        hipLaunchKernelGGL(k_samples<27>, grid, block, 0, s, int(coarse->nx), int(coarse->ny), int(coarse->nz), work,
                           ss, gs, its, feat, target, weight);
    return herr(hipGetLastError());
}

}  // extern "C"

namespace ssmd {
// tau_t = max(floor, scale * network(feat)) over n cells (the LEARNED closure).
int launch_closure_update(const void *packed, const uint16_t *feat, float *tau_t, uint32_t n, float scale,
                          float floor, hipStream_t s) {
    hipLaunchKernelGGL(mlp::k_forward, dim3(mlp::forward_blocks(n)), dim3(256), 0, s,
                       static_cast<const uint8_t *>(packed), feat, tau_t, n, scale, floor, 1);
    return herr(hipGetLastError());
}
}  // namespace ssmd
