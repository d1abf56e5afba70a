// kernels.hpp -- gfx950 kernels of the per-timestep path.
//
// Data layout in HBM (one field set):
//   scalar fields  float[N], idx = x + nx*(y + ny*z)
//   f, f_next      Q planes of N pop elements (f32 / f16 / bf16 deviations)
//   g, g_next      4 planes of N floats: the D3Q7 phase populations in
//                  moment form (g_phi, o_x, o_y, o_z), see phase_column;
//                  usually one buffer (k_phase reads g before k_lattice of
//                  the same step writes it)
// Populations are source-indexed ("pull"): slot [i][s] holds what leaves
// cell s along c_i; the population arriving at cell a along c_i is slot
// [i][a - c_i].  Every f element is read once and written once per step.
//
// Per step:
//   k_phase    phi(x) = sum_k g_k(x - c_k) on interior cells          (K1)
//   k_lattice  fused phase::computeForces + lbm::computeMoments +
//              lbm::streamCollide: LDS-staged phi tile with a radius-2
//              halo marched along z, masked normals -> curvature -> F,
//              regularized collision, f and g written at the own cell  (K2-K4)
//   boundary kernels in reference order (K5-K9, + periodic y / z)
#pragma once

#include "../../include/ssmd.h"
#include "lattice.hpp"

namespace ssmd {

struct JetK {
    int enabled;
    float ca, cb, r2, u, phi;
};

struct KParams {
    int nx, ny, nz;
    uint32_t sxy;  // nx*ny
    uint32_t N;    // cells
    float visc_water, visc_diff;
    float omco_water, omco_oil, omco_ref;
    float sigma, gamma;
    float perturb;
    uint64_t seed;
    JetK water, oil;
    int wrap_x, wrap_y, wrap_z;  // SSMD_WRAP_*
    // planes whose moments a boundary kernel reads on every step: bit 0 y = 1
    // (water inlet), bit 1 y = ny-2 (outflow y), bit 2 z = 1 (oil inlet),
    // bit 3 z = nz-2 (outflow z)
    int bc_planes;
    // the lattice kernel stores g through the caches (k_phase reads it next)
    // when one step's g fits the 256 MB Infinity Cache, else non-temporally
    int g_cached;
    // FP32 D3Q19 lattice kernels pull with non-temporal loads (D3Q27 always
    // does): set where g_cached, so the f stream leaves g and phi cached
    int nt_pull;
    int closure;                 // SSMD_CLOSURE_*
    float smag;                  // 18 Cs^2
    float strain_scale, grad_scale;
    float flux3;                 // LEARNED phase-field flux closure: flux_coeff / 3
};

struct Macro {
    float *rho, *ux, *uy, *uz, *pxx, *pyy, *pzz, *pxy, *pxz, *pyz, *fsx, *fsy, *fsz;
    float *tau_t;    // LEARNED closure: eddy relaxation time (read)
    uint16_t *feat;  // LEARNED closure: 8 bf16 network inputs per cell (written)
};

// The lattice kernel reads its Macro where it is stored (the kernel argument
// segment, or a jet table), through a pointer made opaque at each use: the
// compiler then loads the few pointers a boundary plane needs when it needs
// them instead of holding all 15 in scalar registers across the z loop,
// which would spill other uniform values.
using MacroRef = const __attribute__((address_space(4))) Macro *;
__device__ __forceinline__ MacroRef macro_at(MacroRef m) {
    asm volatile("" : "+s"(m));
    return m;
}
// A Macro read directly, held in registers where the compiler chooses (the
// ensemble kernel's jet table: reading it on demand there measured 18% slower).
__device__ __forceinline__ const Macro *macro_at(const Macro *m) { return m; }
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T *global_ptr(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

__device__ __forceinline__ uint32_t cell_id(const KParams &k, int x, int y, int z) {
    return uint32_t(x) + uint32_t(k.nx) * (uint32_t(y) + uint32_t(k.ny) * uint32_t(z));
}

__device__ __forceinline__ bool is_interior(const KParams &k, int x, int y, int z) {
    return x >= 1 && x <= k.nx - 2 && y >= 1 && y <= k.ny - 2 && z >= 1 && z <= k.nz - 2;
}

// --------------------------------------------------------------------------
// Workgroups are dealt to the 8 XCDs in turn (block b runs on XCD b % 8), so
// block b of a launch of `blocks` blocks takes work unit start(b % 8) + b / 8:
// each XCD works through one contiguous range of units and finds the stencil
// neighbours of a unit (x +- 1 lines, y +- 1 rows, planes z +- 1) in its own
// L2.  A bijection of [0, blocks) for any count.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t blocks) {
    const uint32_t xcd = b & 7u, per = blocks >> 3, extra = blocks & 7u;  // XCD j owns per + (j < extra)
    return xcd * per + (xcd < extra ? xcd : extra) + (b >> 3);
}

// Ghost-column cells (x = 0, nx-1) of interior rows store back the phi they
// hold (boundary data of the previous step) so that every phi line a phase
// kernel writes is written whole.
__device__ __forceinline__ void phase_keep(float *__restrict__ phi, uint32_t c) {
    float v = phi[c];
    asm volatile("" : "+v"(v));  // an opaque value: the store is not elided
    phi[c] = v;
}

// The single-box phase kernel: 64 x 4 tiles of interior rows, each marching
// over kPhaseLZ planes; tiles numbered x-block fastest, then y-block, then
// z-chunk, one block per tile: ceil(nx/64) ceil((ny-2)/4) ceil((nz-2)/kPhaseLZ).
constexpr int kPhaseLZ = 4;
// Planes per block of k_phase_quad: 16 on boxes of up to 2^24 cells, 4 above
// (128 x 128 x 512: 27.5 us at 16 against 29.7 at 4; 256 x 256 x 1024: 276 us
// at 4 against 289 at 16, in isolation).
inline int phase_quad_lz(uint64_t cells) { return cells <= (uint64_t(1) << 24) ? 16 : 4; }

// K1: phi = sum of the seven arriving phase populations (phaseField.cuh:31-52).
//
// The phase populations are stored in moment form, four planes of g:
// g_phi = g[0..N), o = (g[N..2N), g[2N..3N), g[3N..4N)), one value per source
// cell s, with population k leaving s equal to
//     w_k (g_phi(s) + c_k . o(s)),  w_0 = 1/4, w_1..6 = 1/8.
// Every population the collision emits has this form (phase_moments), and the
// boundary kernels write the one population they set as g_phi with o = 0.
// The arriving sum below forms each population with the same two operations
// the collision used (an add, then a power-of-two scale), so phi is the sum of
// the same seven FP32 values, in the same order, as with population storage.
//
// phi of one cell from its seven arriving populations, summed in order:
// rest, +x (source x-1), -x (x+1), +y, -y, +z, -z, each formed as
// w_k (g_phi + c_k . o) of its source cell.
__device__ __forceinline__ float phase_sum(float cg, float g1, float o1, float g2, float o2, float g3, float o3,
                                           float g4, float o4, float bg, float bz, float ag, float az) {
    float s = weight7(0) * cg;
    s += weight7(1) * (g1 + o1);
    s += weight7(2) * (g2 - o2);
    s += weight7(3) * (g3 + o3);
    s += weight7(4) * (g4 - o4);
    s += weight7(5) * (bg + bz);
    s += weight7(6) * (ag - az);
    return s;
}

// One column (x, y), planes z0..z1-1: the +z and -z sources of a plane are
// the own column's values of the planes below and above, carried in
// registers from plane to plane (each plane loads only the one above).
// Source slots across a periodic face are read from their periodic image.
__device__ __forceinline__ void phase_column(const KParams &k, const float *__restrict__ g, float *__restrict__ phi,
                                             int x, int y, int z0, int z1) {
    int sxp = 0, sxm = 0, syp = 0, sym = 0;
    if (k.wrap_x) {
        const int sh = k.wrap_x == SSMD_WRAP_TRUE ? k.nx - 2 : k.nx - 3;
        if (x == 1) sxp = sh;
        if (x == k.nx - 2) sxm = -sh;
    }
    if (k.wrap_y) {
        const int sh = (k.wrap_y == SSMD_WRAP_TRUE ? k.ny - 2 : k.ny - 3) * k.nx;
        if (y == 1) syp = sh;
        if (y == k.ny - 2) sym = -sh;
    }
    const int zsh = k.wrap_z ? (k.wrap_z == SSMD_WRAP_TRUE ? k.nz - 2 : k.nz - 3) * int(k.sxy) : 0;
    const size_t N = k.N;
    const float *__restrict__ gp = g;
    const float *__restrict__ ox = g + N;
    const float *__restrict__ oy = g + 2 * N;
    const float *__restrict__ oz = g + 3 * N;
    const int sxy = int(k.sxy);
    int ci = int(cell_id(k, x, y, z0));
    const uint32_t sb = uint32_t(ci - sxy + (z0 == 1 ? zsh : 0));
    float bg = gp[sb], bz = oz[sb];  // source of the +z population (plane below)
    float cg = gp[ci], cz = oz[ci];  // own column, this plane
    for (int z = z0; z < z1; ++z, ci += sxy) {
        const uint32_t sa = uint32_t(ci + sxy - (z == k.nz - 2 ? zsh : 0));
        const float ag = gp[sa], az = oz[sa];  // source of the -z population (plane above)
        const uint32_t s1 = uint32_t(ci - 1 + sxp), s2 = uint32_t(ci + 1 + sxm);
        const uint32_t s3 = uint32_t(ci - k.nx + syp), s4 = uint32_t(ci + k.nx + sym);
        phi[ci] = phase_sum(cg, gp[s1], ox[s1], gp[s2], ox[s2], gp[s3], oy[s3], gp[s4], oy[s4], bg, bz, ag, az);
        bg = cg;
        bz = cz;
        cg = ag;
        cz = az;
    }
}

#ifdef SSMD_STEP_KERNELS
// One block per phase tile, dealt to the XCDs in contiguous ranges.
// `counter` (graph steps; else null): the device step index, advanced here,
// at the start of a step, for the inlet noise of that step's face kernels.
__global__ void __launch_bounds__(256) k_phase(KParams k, const float *__restrict__ g, float *__restrict__ phi,
                                               uint64_t *counter) {
    if (counter && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *counter += 1;
    const uint32_t t = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t bx = uint32_t(k.nx + 63) >> 6, by = uint32_t(k.ny - 2 + 3) >> 2;
    const uint32_t row = t / bx;
    const int x = int(t - row * bx) * 64 + int(threadIdx.x);
    const int y = 1 + int(row % by) * 4 + int(threadIdx.y);
    const int z0 = 1 + int(row / by) * kPhaseLZ;
    const int z1 = min(z0 + kPhaseLZ, k.nz - 1);
    if (x > k.nx - 1 || y > k.ny - 2) return;
    if (x == 0 || x == k.nx - 1) {
        for (int z = z0; z < z1; ++z) phase_keep(phi, cell_id(k, x, y, z));
        return;
    }
    phase_column(k, g, phi, x, y, z0, z1);
}

// The phase kernel for rows of a multiple of 4 cells (every plane then starts
// on a 16-byte boundary): a lane takes the 4 x-adjacent cells of one 16-byte
// word of each moment plane and a block 256 consecutive such quads of the
// interior rows (x fastest), marching over LZ planes as phase_column does,
// with blocks dealt to the XCDs in contiguous ranges.  Per plane a lane
// loads five 16-byte words (g_phi, o_x of its row; g_phi, o_y of the rows
// y -+ 1; g_phi, o_z of the plane above -- the planes below and its own are
// carried) and the two 4-byte ends x-1, x+4 of the g_phi and o_x rows, and
// stores one 16-byte word of phi: a quarter of the memory instructions of
// k_phase, every cell summed by phase_sum as there.  The loads of plane z+1
// are issued before plane z is summed (software pipelining: 290 -> 276 us on
// 256 x 256 x 1024).  Ghost-column cells (x = 0, nx-1) keep the phi they hold.
// Grid: ceil(quads / 256) ceil((nz - 2) / LZ) blocks of 256.  `counter` as
// for k_phase.
// Quad q (< (nx/4)(ny-2), x fastest) of the interior rows, planes of z-chunk
// zc.  NTS: phi stored non-temporally (boxes whose phi exceeds the Infinity
// Cache: 3.7% off the kernel on 256 x 256 x 1024).
template <int LZ, bool NTS>
__device__ __forceinline__ void phase_quad_column(const KParams &k, const float *__restrict__ g,
                                                  float *__restrict__ phi, uint32_t q, uint32_t zc) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    const uint32_t qx = uint32_t(k.nx) >> 2;
    const uint32_t yq = q / qx;
    const int x = int(q - yq * qx) * 4, y = 1 + int(yq);
    const int z0 = 1 + int(zc) * LZ, z1 = min(z0 + LZ, k.nz - 1);
    const bool first = x == 0, last = x == k.nx - 4;  // quads holding a ghost column
    // periodic x: cell 1 (first quad) pulls +x from the image of x = 0, cell
    // nx-2 (last quad) pulls -x from the image of x = nx-1; element offset wo
    const bool wf = k.wrap_x && first, wl = k.wrap_x && last;
    int wo = 0;
    if (k.wrap_x) {
        const int sh = k.wrap_x == SSMD_WRAP_TRUE ? k.nx - 2 : k.nx - 3;
        wo = first ? sh : 3 - sh;
    }
    int syp = 0, sym = 0;
    if (k.wrap_y) {
        const int sh = (k.wrap_y == SSMD_WRAP_TRUE ? k.ny - 2 : k.ny - 3) * k.nx;
        if (y == 1) syp = sh;
        if (y == k.ny - 2) sym = -sh;
    }
    const int zsh = k.wrap_z ? (k.wrap_z == SSMD_WRAP_TRUE ? k.nz - 2 : k.nz - 3) * int(k.sxy) : 0;
    const size_t N = k.N;
    const float *__restrict__ gp = g;
    const float *__restrict__ ox = g + N;
    const float *__restrict__ oy = g + 2 * N;
    const float *__restrict__ oz = g + 3 * N;
    const int sxy = int(k.sxy);
    auto ld4 = [](const float *base, int i) { return *reinterpret_cast<const F4 *>(base + i); };
    const int r3o = -k.nx + syp, r4o = k.nx + sym;  // rows y - 1, y + 1 (periodic images)
    // the loads of one plane: sources above (g_phi, o_z), own o_x, the x row
    // ends, the wrap image and the y -+ 1 rows
    struct Plane {
        F4 ag, az, xo, g3, o3, g4, o4;
        float gm, om, gq, oq, gw, ow;
    };
    auto load = [&](int z, int c) {
        Plane l;
        const int sa = c + sxy - (z == k.nz - 2 ? zsh : 0);
        l.ag = ld4(gp, sa);
        l.az = ld4(oz, sa);
        l.xo = ld4(ox, c);
        l.gm = gp[c - 1];
        l.om = ox[c - 1];
        l.gq = gp[c + 4];
        l.oq = ox[c + 4];
        l.gw = 0.f;
        l.ow = 0.f;
        if (wf || wl) {
            l.gw = gp[c + wo];
            l.ow = ox[c + wo];
        }
        l.g3 = ld4(gp, c + r3o);
        l.o3 = ld4(oy, c + r3o);
        l.g4 = ld4(gp, c + r4o);
        l.o4 = ld4(oy, c + r4o);
        return l;
    };
    int ci = int(cell_id(k, x, y, z0));
    const int sb = ci - sxy + (z0 == 1 ? zsh : 0);
    F4 bg = ld4(gp, sb), bz = ld4(oz, sb);  // sources of the +z populations (plane below)
    F4 cg = ld4(gp, ci), cz = ld4(oz, ci);  // own cells, this plane
    Plane cur = load(z0, ci);
    for (int z = z0; z < z1; ++z, ci += sxy) {
        Plane nxt = cur;
        if (z + 1 < z1) nxt = load(z + 1, ci + sxy);
        // +x source of cell e: element e-1 of the row; -x source: element e+1
        const float G1[4] = {cur.gm, wf ? cur.gw : cg.x, cg.y, cg.z};
        const float O1[4] = {cur.om, wf ? cur.ow : cur.xo.x, cur.xo.y, cur.xo.z};
        const float G2[4] = {cg.y, cg.z, wl ? cur.gw : cg.w, cur.gq};
        const float O2[4] = {cur.xo.y, cur.xo.z, wl ? cur.ow : cur.xo.w, cur.oq};
        F4 out;
#pragma unroll
        for (int e = 0; e < 4; ++e)
            out[e] = phase_sum(cg[e], G1[e], O1[e], G2[e], O2[e], cur.g3[e], cur.o3[e], cur.g4[e], cur.o4[e], bg[e],
                               bz[e], cur.ag[e], cur.az[e]);
        if (first || last) {  // the ghost column's phi is boundary data: kept
            const F4 old = ld4(phi, ci);
            if (first) out.x = old.x;
            if (last) out.w = old.w;
        }
        if constexpr (NTS)
            __builtin_nontemporal_store(out, reinterpret_cast<F4 *>(phi + ci));
        else
            *reinterpret_cast<F4 *>(phi + ci) = out;
        bg = cg;
        bz = cz;
        cg = cur.ag;
        cz = cur.az;
        cur = nxt;
    }
}
template <int LZ>
__global__ void __launch_bounds__(256) k_phase_quad(KParams k, const float *__restrict__ g,
                                                    float *__restrict__ phi, uint64_t *counter) {
    if (counter && blockIdx.x == 0 && threadIdx.x == 0) *counter += 1;
    const uint32_t rowq = (uint32_t(k.nx) >> 2) * uint32_t(k.ny - 2), pb = (rowq + 255u) >> 8;
    const uint32_t t = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t zc = t / pb, q = (t - zc * pb) * 256u + threadIdx.x;
    if (q >= rowq) return;
    phase_quad_column<LZ, LZ == 4>(k, g, phi, q, zc);  // 4-plane chunks: boxes past 2^24 cells
}
#endif  // SSMD_STEP_KERNELS


// --------------------------------------------------------------------------
// Fused lattice kernel.  A block of TX x TY threads owns an x-y tile and
// marches over a chunk of z planes.  LDS rings hold 5 planes of phi-hat
// (phi on interior cells, 0 on the ghost layer and outside; phaseField.cuh:
// 125-142) with a 2-cell halo and 4 planes of the masked normals with a
// 1-cell halo: every thread gets the radius-2 neighbourhood of the force.
// The fourth normal plane lets one barrier per z plane order the ring
// writes of plane z+1 against the reads of plane z.
template <int TX, int TY>
struct LatticeSmem {
    static constexpr int HX = TX + 4, HY = TY + 4, HP = HX * HY;
    static constexpr int NXl = TX + 2, NYl = TY + 2, NP = NXl * NYl;
    float m[5][HP];     // phi-hat ring
    float n[4][4][NP];  // masked normal x, y, z and |grad phi-hat| (planes x component)
};

// Byte offset of the periodic image of a ghost source slot, for the
// directions that pull across a periodic face (the reference's periodic
// copies, BoundaryConditions.cuh:305-332, and the TRUE wrap, folded into the
// addresses: the copy kernels never run).  `lo` = cell index 1 (its source
// slot 0 is a ghost for c > 0), `hi` = cell n-2 (source n-1 for c < 0).
__device__ __forceinline__ int wrap_shift(int mode, int n) { return mode == SSMD_WRAP_TRUE ? n - 2 : n - 3; }

// Correctly rounded sqrt/division are not needed here: one-ulp hardware ops.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Relaxation time of a closure other than NONE (no reference counterpart).
// tau0 = 1/2 + 3 nu(phi); `grad` = |grad phi-hat| of the cell.
//   SMAGORINSKY: Hou et al. (1996), tau = (tau0 + sqrt(tau0^2 + 18 Cs^2
//                sqrt(2 Pi:Pi) / rho)) / 2 with Pi the non-equilibrium moments.
//   LEARNED:     tau = tau0 + tau_t; writes the network inputs of the cell:
//                the strain rate S = -3 Pi / (2 rho tau) (scaled), phi and
//                the scaled |grad phi-hat|, as 8 bf16 in one 16-byte store.
__device__ __forceinline__ uint32_t bf16_pair(float a, float b) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));  // one v_cvt_pk_bf16_f32
}
template <class MP>
__device__ __forceinline__ float closure_tau(const KParams &k, MP macp, const Cell &m, uint32_t c, float ph,
                                             float grad, float tau0, bool write_inputs, float &tau_t) {
    tau_t = 0.f;
    if (k.closure == SSMD_CLOSURE_SMAGORINSKY) {
        const float q = m.pxx * m.pxx + m.pyy * m.pyy + m.pzz * m.pzz +
                        2.0f * (m.pxy * m.pxy + m.pxz * m.pxz + m.pyz * m.pyz);
        return 0.5f * (tau0 + __builtin_amdgcn_sqrtf(tau0 * tau0 + k.smag * __builtin_amdgcn_sqrtf(2.0f * q) *
                                                                         __builtin_amdgcn_rcpf(m.rho)));
    }
    const auto mac = macro_at(macp);
    tau_t = global_ptr(mac->tau_t)[c];
    const float tau = tau0 + tau_t;
    const float s = -1.5f * k.strain_scale * __builtin_amdgcn_rcpf(m.rho * tau);
    uint4 w;
    w.x = bf16_pair(m.pxx * s, m.pyy * s);
    w.y = bf16_pair(m.pzz * s, m.pxy * s);
    w.z = bf16_pair(m.pxz * s, m.pyz * s);
    w.w = bf16_pair(ph, grad * k.grad_scale);
    // non-temporal: the network pass reads these once, and cached stores of
    // them would push g and phi out of the Infinity Cache (+0.6% on the step)
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    if (write_inputs)
        __builtin_nontemporal_store(U4{w.x, w.y, w.z, w.w}, reinterpret_cast<U4 *>(mac->feat + size_t(c) * 8));
    return tau;
}

// Surface-tension force of one cell from the masked-normal ring
// (phaseField.cuh:246-293): curvature 3 sum_i w_i c_i . n(x + c_i) over
// opposite pairs grouped by weight class, F = -sigma kappa |grad phi-hat| n.
// nzm / nz0 / nzp are the ring slots of planes z-1, z, z+1; ncen the cell's
// index in a ring plane of row length NXl.  nc = (n, |grad phi-hat|) of the cell.
template <int Q, int NXl, int NP>
__device__ __forceinline__ void surface_force(const float (&s_n)[4][4][NP], int nzm, int nz0, int nzp, int ncen,
                                              float sigma, float4 &nc, float &kx, float &ky, float &kz) {
    nc = make_float4(s_n[nz0][0][ncen], s_n[nz0][1][ncen], s_n[nz0][2][ncen], s_n[nz0][3][ncen]);
    // component a of the masked normal at offset (dx, dy, dz), addressed from
    // the (-1, -1) corner of each plane so that every offset is a non-negative
    // immediate of one LDS address per plane
    const float *cm = &s_n[nzm][0][ncen - NXl - 1], *c0 = &s_n[nz0][0][ncen - NXl - 1],
                *cp = &s_n[nzp][0][ncen - NXl - 1];
    auto nat = [&](int a, int dx, int dy, int dz) {
        const float *b = dz < 0 ? cm : (dz > 0 ? cp : c0);
        return b[a * NP + (dy + 1) * NXl + dx + 1];
    };
    // the curvature needs only the sum over axes: one sum per weight class
    Sum acc[3];
    sfor<0, (Q - 1) / 2>([&](auto Pp) {
        constexpr int i = 2 * Pp + 1;
        constexpr Dir d = kDir[i];
        constexpr int cls = i <= 6 ? 0 : (i <= 18 ? 1 : 2);
        if constexpr (d.x != 0) {
            const float a = nat(0, d.x, d.y, d.z), b = nat(0, -d.x, -d.y, -d.z);
            acc[cls].add(d.x > 0 ? a - b : b - a);
        }
        if constexpr (d.y != 0) {
            const float a = nat(1, d.x, d.y, d.z), b = nat(1, -d.x, -d.y, -d.z);
            acc[cls].add(d.y > 0 ? a - b : b - a);
        }
        if constexpr (d.z != 0) {
            const float a = nat(2, d.x, d.y, d.z), b = nat(2, -d.x, -d.y, -d.z);
            acc[cls].add(d.z > 0 ? a - b : b - a);
        }
    });
    // 3 w_class, the prefactor folded into the weights
    constexpr float w1 = 3.0f * weight<Q>(1), w2 = 3.0f * weight<Q>(7), w3 = Q == 27 ? 3.0f * weight<Q>(19) : 0.f;
    float curvature = w1 * acc[0].v + w2 * acc[1].v;
    if constexpr (Q == 27) curvature += w3 * acc[2].v;
    const float st = -sigma * curvature * nc.w;
    kx = st * nc.x;
    ky = st * nc.y;
    kz = st * nc.z;
}

// Normal of the stored phi for the phase update (lbm.cu:157-224).  It equals
// the masked normal nc unless the stencil reaches the ghost layer; there the
// ghost values are added back, face by face (a neighbour on two ghost faces
// is taken by the first).  Only for interior cells.
template <int Q>
__device__ __forceinline__ void stored_phi_normal(const KParams &k, const float *__restrict__ phi, int x, int y,
                                                  int z, bool active, const float4 &nc, float &nx, float &ny,
                                                  float &nz, float &gs) {
    nx = nc.x;
    ny = nc.y;
    nz = nc.z;
    gs = nc.w;
    const bool fx_lo = x == 1, fx_hi = x == k.nx - 2, fy_lo = y == 1, fy_hi = y == k.ny - 2;
    const bool fz_lo = z == 1, fz_hi = z == k.nz - 2;
    if (!(active && (fx_lo || fx_hi || fy_lo || fy_hi || fz_lo || fz_hi))) return;
    const float g = nc.w + 1e-9f;
    float gx = nc.x * g, gy = nc.y * g, gz = nc.z * g;
    // add 3 w_i c_i phi(x + c_i) for every neighbour on the ghost layer
    auto face_sum = [&](auto axis_c, auto sign_c, bool on) {
        constexpr int AX = decltype(axis_c)::value, SG = decltype(sign_c)::value;
        if (!on) return;
        sfor<1, Q>([&](auto I) {
            constexpr int i = I;
            constexpr Dir d = kDir[i];
            constexpr int cv[3] = {d.x, d.y, d.z};
            if constexpr (cv[AX] == SG) {
                bool taken = false;
                if constexpr (AX >= 1 && d.x != 0) taken = taken || (d.x < 0 ? fx_lo : fx_hi);
                if constexpr (AX >= 2 && d.y != 0) taken = taken || (d.y < 0 ? fy_lo : fy_hi);
                const float v = 3.0f * weight<Q>(i) * phi[cell_id(k, x + d.x, y + d.y, z + d.z)];
                const float vv = taken ? 0.f : v;
                if constexpr (d.x != 0) gx += float(d.x) * vv;
                if constexpr (d.y != 0) gy += float(d.y) * vv;
                if constexpr (d.z != 0) gz += float(d.z) * vv;
            }
        });
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using Pm = std::integral_constant<int, -1>;
    using Pp = std::integral_constant<int, 1>;
    face_sum(I0{}, Pm{}, fx_lo);
    face_sum(I0{}, Pp{}, fx_hi);
    face_sum(I1{}, Pm{}, fy_lo);
    face_sum(I1{}, Pp{}, fy_hi);
    face_sum(I2{}, Pm{}, fz_lo);
    face_sum(I2{}, Pp{}, fz_hi);
    const float ind = fast_sqrt(gx * gx + gy * gy + gz * gz);
    const float inv = fast_rcp(ind + 1e-9f);
    nx = gx * inv;
    ny = gy * inv;
    nz = gz * inv;
    gs = ind;
}

// Bytes of global memory, for pointers the compiler must not lose the
// address space of when they pass through an opaque register copy.
using GChar = __attribute__((address_space(1))) char;

// Stores the Q post-collision populations of one cell, emitted in the order
// 0, then opposite pairs (2k-1, 2k), each at its own plane (plane i at
// plane0 + i * SB, advanced by one scalar add per plane) and byte offset
// `so`, non-temporally: the next step reads them from HBM.  16-bit formats
// round each pair with one packed conversion and store its halves.
template <class P, class Collide>
__device__ __forceinline__ void store_pops(uint32_t so, GChar *plane, size_t SB, Collide &&collide_cell) {
    using PT = typename P::T;
    using AS1 = __attribute__((address_space(1))) PT;
    if constexpr (sizeof(PT) == 2) {
        float first = 0.f;
        GChar *first_plane = plane;
        collide_cell([&](int i, float v) {
            if (i > 0) plane += SB;
            asm volatile("" : "+s"(plane));
            if (i == 0) {
                __builtin_nontemporal_store(P::store(v), reinterpret_cast<AS1 *>(plane + so));
            } else if (i % 2 == 1) {
                first = v;
                first_plane = plane;
            } else {
                const uint32_t w = Pack2<P>::pack(first, v);
                __builtin_nontemporal_store(PT(w), reinterpret_cast<AS1 *>(first_plane + so));
                __builtin_nontemporal_store(PT(w >> 16), reinterpret_cast<AS1 *>(plane + so));
            }
        });
    } else {
        collide_cell([&](int i, float v) {
            if (i > 0) plane += SB;
            asm volatile("" : "+s"(plane));
            __builtin_nontemporal_store(P::store(v), reinterpret_cast<AS1 *>(plane + so));
        });
    }
}

// Planes whose populations are in flight while a plane collides.  A 16-bit
// pull moves half the bytes per instruction of an FP32 one; on D3Q27 a
// second plane in flight restores them (+14% for the 16-bit kernels, still
// 2 waves per SIMD).  D3Q19 (-6% f16, -1.5% f32) and D3Q27 FP32 (-1.5%, 3 -> 2
// waves per SIMD) keep one.
template <int Q, class P> constexpr int pull_depth() { return Q == 27 && sizeof(typename P::T) == 2 ? 2 : 1; }


// One tile: cells x0..x0+TX-1, y0..y0+TY-1, planes zb..ze-1 (threadIdx = (tx, ty)).
// TX is a multiple of 64, so a wavefront covers part of one x row: y is
// uniform across a wavefront and z across the block.
template <int Q, class P, int TX, int TY, bool MACROS, class MP>
__device__ __forceinline__ void lattice_tile(const KParams &k, const typename P::T *__restrict__ fin,
                                             typename P::T *__restrict__ fout, float *__restrict__ gout,
                                             const float *__restrict__ phi, MP macp, int x0, int y0,
                                             int zb, int ze, LatticeSmem<TX, TY> &sm,
                                             int thread_x = int(threadIdx.x), int thread_y = int(threadIdx.y)) {
    static_assert(TX % 64 == 0, "a wavefront must not span two rows");
    using S = LatticeSmem<TX, TY>;
    using PT = typename P::T;
    constexpr int HX = S::HX, HP = S::HP, NXl = S::NXl, NP = S::NP;
    constexpr int NT = TX * TY;
    constexpr int EPT = (HP + NT - 1) / NT;  // phi-plane elements per thread
    auto &s_m = sm.m;
    auto &s_n = sm.n;

    const int tx = thread_x, tid = tx + TX * thread_y;
    // a wavefront lies in one row: make that visible to the compiler
    const int ty = __builtin_amdgcn_readfirstlane(thread_y);
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x >= 1 && x <= k.nx - 2 && y >= 1 && y <= k.ny - 2;
    // The ghost-column cells x = 0 and x = nx-1 of interior rows are never
    // inlet cells (inlets lie on the y = 0 and z = 0 faces) and no kernel
    // writes their population slots, so those hold the t = 0 equilibrium at
    // rho 1, u 0, phi 0 -- exactly 0 -- for the whole run.  Their lanes store
    // that 0 in the same instructions as the interior lanes: every population
    // cache line is then written whole, which HBM takes ~20% faster than
    // lines with their first or last element masked off.
    const bool ghost_x = (x == 0 || x == k.nx - 1) && y >= 1 && y <= k.ny - 2;
    const bool lane = active || ghost_x;
    const size_t N0 = k.N;
    const uint32_t sxy = k.sxy;

    // per-lane x wrap of the pulled source slot (bytes), uniform y wrap (cells)
    int xsp = 0, xsm = 0;
    if (k.wrap_x) {
        const int sh = wrap_shift(k.wrap_x, k.nx);
        if (x == 1) xsp = sh;
        if (x == k.nx - 2) xsm = -sh;
    }
    // Ghost-column lanes pull only to keep the wave's loads uniform (their
    // results are discarded): cancel the x step so they read their own
    // column, on lines the interior lanes' c_x = 0 pulls fetch anyway.
    if (ghost_x) {
        xsp = 1;
        xsm = -1;
    }
    long ysp = 0, ysm = 0;
    if (k.wrap_y) {
        const long sh = long(wrap_shift(k.wrap_y, k.ny)) * k.nx;
        if (y == 1) ysp = sh;
        if (y == k.ny - 2) ysm = -sh;
    }

    // phi plane p: element e = tid + q NT of the (TX+4)(TY+4) halo plane.
    // Its in-plane byte offset is the same on every plane; ~0u marks an
    // element outside the halo plane or on the ghost layer (read as 0).
    uint32_t hoff[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = tid + q * NT;
        const int gx = x0 + e % HX - 2, gy = y0 + e / HX - 2;
        const bool ii = e < HP && gx >= 1 && gx <= k.nx - 2 && gy >= 1 && gy <= k.ny - 2;
        hoff[q] = ii ? (uint32_t(gx) + uint32_t(k.nx) * uint32_t(gy)) * 4u : ~0u;
    }
    auto fetch_plane = [&](int p, float (&r)[EPT]) {
        const bool plane = p >= 1 && p <= k.nz - 2;
        // (a plane outside the grid reads nothing: its base is never formed)
        const char *base = reinterpret_cast<const char *>(phi + size_t(plane ? p : 0) * k.sxy);
#pragma unroll
        for (int q = 0; q < EPT; ++q)
            r[q] = plane && hoff[q] != ~0u ? *reinterpret_cast<const float *>(base + hoff[q]) : 0.f;
    };
    auto store_plane = [&](int p, const float (&r)[EPT]) {
        const int sl = (p + 5) % 5;
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * NT;
            if (e < HP) s_m[sl][e] = r[q];
        }
    };
    // normal-ring elements e = tid + q NT of this thread: the phi-hat index
    // of their (-1, -1) stencil corner, the same on every plane
    constexpr int NPT = (NP + NT - 1) / NT;
    int ncorner[NPT];
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int e = tid + q * NT;
        ncorner[q] = (e / NXl) * HX + e % NXl;  // (hy - 1) HX + hx - 1 with hx, hy = e % NXl + 1, e / NXl + 1
    }
    auto normals = [&](int p) {
        const int sl = p % 4;
        const int pm = (p + 4) % 5, p0 = (p + 5) % 5, pp = (p + 6) % 5;
#pragma unroll
        for (int q = 0; q < NPT; ++q) {
            const int e = tid + q * NT;
            if (q == NPT - 1 && e >= NP) break;
            const float *cm = &s_m[pm][ncorner[q]], *c0 = &s_m[p0][ncorner[q]], *cp = &s_m[pp][ncorner[q]];
            auto at = [&](int dx, int dy, int dz) {
                const float *b = dz < 0 ? cm : (dz > 0 ? cp : c0);
                return b[(dy + 1) * HX + dx + 1];
            };
            float gx, gy, gz;
            iso_grad<Q>(at, gx, gy, gz);
            const float ind = fast_sqrt(gx * gx + gy * gy + gz * gz);
            const float inv = fast_rcp(ind + 1e-9f);
            s_n[sl][0][e] = gx * inv;
            s_n[sl][1][e] = gy * inv;
            s_n[sl][2][e] = gz * inv;
            s_n[sl][3][e] = ind;
        }
    };

    {
        float r[EPT];
        for (int p = zb - 2; p <= zb + 2; ++p) {
            fetch_plane(p, r);
            store_plane(p, r);
        }
    }
    __syncthreads();
    normals(zb - 1);
    normals(zb);
    float nxt[EPT];
    if (zb + 1 < ze) fetch_plane(zb + 3, nxt);  // one plane ahead, in registers
    __syncthreads();  // plane zb-2 (read above) is the first ring slot reused

    // Pull the Q populations arriving at (x, y, zz), as stored.  Each pull's
    // uniform base is computed on the scalar unit right where it is used
    // (not held, and spilled, across the loop).
    auto pull_as = [&]<bool NT>(int zz, PT (&f)[Q]) {
        const uint32_t c = cell_id(k, x, y, zz);
        long zsp = 0, zsm = 0;
        if (k.wrap_z) {
            const long sh = long(wrap_shift(k.wrap_z, k.nz)) * long(sxy);
            if (zz == 1) zsp = sh;
            if (zz == k.nz - 2) zsm = -sh;
        }
        // uniform byte steps from a population plane to the source row of
        // a c_y = +-1 or c_z = +-1 population, periodic wraps included
        constexpr long B = long(sizeof(PT));
        const long yp = (ysp - long(k.nx)) * B, ym = (ysm + long(k.nx)) * B;
        const long zp = (zsp - long(sxy)) * B, zm = (zsm + long(sxy)) * B;
        // uniform base per direction + one of three shared 32-bit byte offsets
        const uint32_t o0 = c * uint32_t(sizeof(PT));
        const uint32_t op = uint32_t(int(c) + xsp) * uint32_t(sizeof(PT));  // c_x > 0 sources
        const uint32_t om = uint32_t(int(c) + xsm) * uint32_t(sizeof(PT));  // c_x < 0 sources
        // plane i's base is the previous one's plus N: one 64-bit scalar add
        // per direction, kept from being folded back into i * N
        const GChar *plane = reinterpret_cast<const GChar *>(reinterpret_cast<uintptr_t>(fin));
        const size_t NB = size_t(N0) * sizeof(PT);
        sfor<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr Dir d = kDir[i];
            if constexpr (i > 0) plane += NB;
            asm volatile("" : "+s"(plane));
            long shift = -long(d.x) * B;
            if constexpr (d.y > 0) shift += yp;
            if constexpr (d.y < 0) shift += ym;
            if constexpr (d.z > 0) shift += zp;
            if constexpr (d.z < 0) shift += zm;
            const GChar *base = plane + shift;
            const uint32_t off = d.x > 0 ? op : (d.x < 0 ? om : o0);
            const auto *src = reinterpret_cast<const __attribute__((address_space(1))) PT *>(base + off);
            if constexpr (NT)
                f[i] = __builtin_nontemporal_load(src);
            else
                f[i] = *src;
        });
    };
    // Non-temporal pulls (measured): D3Q27 reads faster with them on every
    // box.  FP32 D3Q19 where one step's g fits the Infinity Cache
    // (k.nt_pull): they leave g and phi there for the next phase update and
    // lattice kernel (-7% step on 128x256x256), elsewhere they are slower
    // (-4.7% on 256x256x1024).  16-bit D3Q19 pulls slower with them on every
    // box, and even an untaken branch to them costs it 3%: never.
    constexpr bool kNtAlways = Q == 27, kNtMaybe = Q != 27 && sizeof(PT) == 4;
    auto pull = [&](int zz, PT (&f)[Q]) {
        if constexpr (kNtAlways) {
            pull_as.template operator()<true>(zz, f);
        } else if constexpr (kNtMaybe) {
            if (k.nt_pull)
                pull_as.template operator()<true>(zz, f);
            else
                pull_as.template operator()<false>(zz, f);
        } else {
            pull_as.template operator()<false>(zz, f);
        }
    };
    // The populations of the next D planes are in flight while plane z
    // collides: the regularized collision needs only the moments, so a
    // register set is dead once they are formed and takes plane z+D.
    constexpr int D = pull_depth<Q, P>();
    PT fa[Q], fb[D == 2 ? Q : 1];
    if (lane) {
        pull(zb, fa);
        if constexpr (D == 2)
            if (zb + 1 < ze) pull(zb + 1, fb);
    }

    auto step = [&](int z, PT (&fr)[Q]) {
        // (1) normals of plane z+1 (phi-hat planes z..z+2) into ring slot
        //     (z+1)%4; then the prefetched phi plane z+3 into the slot of
        //     plane z-2 and the fetch of z+4.  Before the barrier a wave may
        //     still be in (2) of plane z-1, which reads normal slots z-2..z
        //     and phi slot z-1: none of these is written here.
        normals(z + 1);
        if (z + 1 < ze) {
            store_plane(z + 3, nxt);
            if (z + 2 < ze) fetch_plane(z + 4, nxt);
        }
        const uint32_t c = cell_id(k, x, y, z);
        __syncthreads();
        if (!lane) return;

        // (3) surface-tension force from the masked normals (phaseField.cuh:246-293)
        // (2) everything below reads normal slots z-1..z+1 and phi slot z
        const int ncen = (ty + 1) * NXl + tx + 1;
        float kx, ky, kz;
        float4 nc;
        surface_force<Q, NXl>(s_n, (z + 3) % 4, z % 4, (z + 1) % 4, ncen, k.sigma, nc, kx, ky, kz);

        // (4) normal of the stored phi for the phase update (interior lanes
        //     only: a ghost-column lane's x -+ 1 leaves the row)
        float nx, ny, nz, gs;
        const float ph = s_m[z % 5][(ty + 2) * HX + tx + 2];
        stored_phi_normal<Q>(k, phi, x, y, z, active, nc, nx, ny, nz, gs);
        const bool fy_lo = y == 1, fy_hi = y == k.ny - 2, fz_lo = z == 1, fz_hi = z == k.nz - 2;

        // (5) moments; then the pull of plane z+D; then the collision.  A
        //     ghost-column lane collides the rest state without force, which
        //     emits exactly 0 for every population and phase moment.
        if (!active) kx = ky = kz = 0.f;
        const Cell m = rest_or(active, moments<Q, P>(fr, kx, ky, kz));
        // D3Q19 16-bit: the pull is unconditional (the last plane of a chunk
        // re-reads plane ze-1 from the caches and discards it), so the register
        // set is rewritten on every plane and the loop carries no copies of it
        // (-14 VALU per wave and plane, 111 -> 103 VGPRs: 2% on the kernel).
        // The other lattices run slower so (D3Q27 FP32 -6%): conditional.
        if constexpr (Q == 19 && sizeof(PT) == 2)
            pull(min(z + D, ze - 1), fr);
        else if (z + D < ze)
            pull(z + D, fr);
        const float nu = __builtin_fmaf(ph, k.visc_diff, k.visc_water);
        float tau_t = 0.f;
        const float omco = 1.0f - fast_rcp(k.closure == SSMD_CLOSURE_NONE
                                               ? 0.5f + 3.0f * nu
                                               : closure_tau(k, macp, m, c, ph, nc.w, 0.5f + 3.0f * nu, active, tau_t));
        // uniform plane base (the previous one's plus N, as in the pull) +
        // shared 32-bit byte offset for every store; non-temporal: the next
        // step reads these from HBM, nothing here does
        const uint32_t so = c * uint32_t(sizeof(PT));
        GChar *splane = reinterpret_cast<GChar *>(reinterpret_cast<uintptr_t>(fout));
        const size_t SB = size_t(N0) * sizeof(PT);
        store_pops<P>(so, splane, SB, [&](auto &&emit) { collide<Q>(m, omco, kx, ky, kz, emit); });
        float sharp = k.gamma * ph * (1.0f - ph);
        // LEARNED flux closure: J = -D_t grad(phi), D_t = max(flux_coeff tau_t / 3, -0.1),
        // enters the emitted first moments 4 J as -4 D_t |grad phi| along the normal
        if (k.closure == SSMD_CLOSURE_LEARNED && k.flux3 != 0.f && active)
            sharp -= 4.0f * fmaxf(k.flux3 * tau_t, -0.1f) * gs;
        const uint32_t go = c * 4u;
        GChar *gplane = reinterpret_cast<GChar *>(reinterpret_cast<uintptr_t>(gout));
        const size_t GB = size_t(N0) * 4u;
        phase_moments(ph, sharp, m.ux, m.uy, m.uz, nx, ny, nz, [&](int kk, float v) {
            if (kk > 0) gplane += GB;
            asm volatile("" : "+s"(gplane));
            auto *dst = reinterpret_cast<__attribute__((address_space(1))) float *>(gplane + go);
            if (k.g_cached)
                *dst = v;
            else
                __builtin_nontemporal_store(v, dst);
        });

        // rho, u, Pi on the planes the inlet / outflow kernels read them from
        // (BoundaryConditions.cuh:90-92, 159-161, 224-226, 284-286); periodic
        // faces read u only on macro steps, which write every moment
        const bool bc_plane = ((k.bc_planes & 1) && fy_lo) || ((k.bc_planes & 2) && fy_hi) ||
                              ((k.bc_planes & 4) && fz_lo) || ((k.bc_planes & 8) && fz_hi);
        if (!active) return;  // ghost-column lanes write populations only
        if (MACROS || bc_plane) {
            const auto mac = macro_at(macp);
            global_ptr(mac->rho)[c] = m.rho;
            global_ptr(mac->ux)[c] = m.ux;
            global_ptr(mac->uy)[c] = m.uy;
            global_ptr(mac->uz)[c] = m.uz;
            global_ptr(mac->pxx)[c] = m.pxx;
            global_ptr(mac->pyy)[c] = m.pyy;
            global_ptr(mac->pzz)[c] = m.pzz;
            global_ptr(mac->pxy)[c] = m.pxy;
            global_ptr(mac->pxz)[c] = m.pxz;
            global_ptr(mac->pyz)[c] = m.pyz;
            if (MACROS) {
                global_ptr(mac->fsx)[c] = kx;
                global_ptr(mac->fsy)[c] = ky;
                global_ptr(mac->fsz)[c] = kz;
            }
        }
    };
    if constexpr (D == 2) {
        for (int z = zb; z < ze; z += 2) {
            step(z, fa);
            if (z + 1 < ze) step(z + 1, fb);
        }
    } else {
        for (int z = zb; z < ze; ++z) step(z, fa);
    }
}

// Arguments of one lattice launch, passed as the kernel's only argument so
// that it sits at offset 0 of the kernel argument segment.
struct LatticeArgs {
    KParams k;
    const void *fin;
    void *fout;
    float *gout;
    const float *phi;
    Macro mac;
    int lz;
    // 0: block (bx, by, bz) takes tile (bx, by, bz), so consecutive tiles of
    // a row land on different XCDs (blocks are dealt to the 8 XCDs in turn);
    // 1: the tiles, numbered x fastest, then y, then z, are dealt to the XCDs
    // in contiguous ranges, so x- and y-neighbour tiles -- and the cache lines
    // their misaligned pulls share -- sit behind one L2
    int xcd_order;
};

template <int Q, class P, int TX, int TY, bool MACROS>
__global__ void __launch_bounds__(TX *TY) k_lattice(LatticeArgs a) {
    __shared__ LatticeSmem<TX, TY> sm;
    using PT = typename P::T;
    const auto *seg = reinterpret_cast<const __attribute__((address_space(4))) LatticeArgs *>(
        reinterpret_cast<uintptr_t>(__builtin_amdgcn_kernarg_segment_ptr()));
    uint32_t bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (a.xcd_order) {
        const uint32_t gxy = gridDim.x * gridDim.y;
        const uint32_t t = xcd_contiguous(bx + gridDim.x * (by + gridDim.y * bz), gxy * gridDim.z);
        bz = t / gxy;
        by = (t - bz * gxy) / gridDim.x;
        bx = t - bz * gxy - by * gridDim.x;
    }
    const int zb = 1 + int(bz) * a.lz;
    lattice_tile<Q, P, TX, TY, MACROS>(a.k, static_cast<const PT *>(a.fin), static_cast<PT *>(a.fout), a.gout, a.phi,
                                       &seg->mac, int(bx) * TX, int(by) * TY, zb, min(zb + a.lz, a.k.nz - 1), sm);
}

// --------------------------------------------------------------------------
// Ensembles: one launch covers every jet.  Work units (tiles, cell blocks,
// face blocks) of jet j occupy the global id range [begin_j, begin_{j+1});
// a block looks its jet up in the table (uniform, <= 64 entries).
struct JetSlot {
    KParams k;
    const void *f;  // populations read by the step
    void *f_next;
    const float *g;
    float *g_next;
    float *phi;
    Macro mac;
    int tiles_x, tiles_y, lz;  // lattice tiling of this jet
    int narrow;                // its tiles are 64 x 8 in a launch of 128 x 4 blocks (k_lattice_batch)
    uint32_t tile_begin;       // first lattice tile id
    uint32_t cell_block_begin; // first block id of the phase kernel (k_phase_batch)
    uint32_t face_begin[4];    // first face block id for faces normal to x, y, z and the x-y ring (64x4 blocks)
};

__device__ __forceinline__ int find_slot(const JetSlot *jets, int njets, uint32_t id, int which) {
    int j = 0;
    while (j + 1 < njets) {
        const JetSlot &n = jets[j + 1];
        const uint32_t b = which == 0 ? n.tile_begin : (which == 1 ? n.cell_block_begin : n.face_begin[which - 2]);
        if (b > id) break;
        ++j;
    }
    return j;
}

#ifdef SSMD_STEP_KERNELS
// Derived statistics of one step (reference derivedFields/operators/:
// timeAverage.cuh:37-82, reynoldsMoments.cuh:37-96, vorticityFields.cuh:37-105).
struct Derived {
    float *avg[4];   // phi, ux, uy, uz
    float *rey[6];   // uxux uyuy uzuz uxuy uxuz uyuz
    float *vort[4];  // x, y, z, magnitude
};

__global__ void __launch_bounds__(256) k_derived(KParams k, const float *__restrict__ phi,
                                                 const float *__restrict__ ux, const float *__restrict__ uy,
                                                 const float *__restrict__ uz, Derived s, float t) {
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= k.N) return;
    auto upd = [t](float *a, uint32_t i, float v) { a[i] = a[i] + (v - a[i]) / t; };
    const float vx = ux[c], vy = uy[c], vz = uz[c];
    if (s.avg[0]) upd(s.avg[0], c, phi[c]);
    if (s.avg[1]) upd(s.avg[1], c, vx);
    if (s.avg[2]) upd(s.avg[2], c, vy);
    if (s.avg[3]) upd(s.avg[3], c, vz);
    if (s.rey[0]) upd(s.rey[0], c, vx * vx);
    if (s.rey[1]) upd(s.rey[1], c, vy * vy);
    if (s.rey[2]) upd(s.rey[2], c, vz * vz);
    if (s.rey[3]) upd(s.rey[3], c, vx * vy);
    if (s.rey[4]) upd(s.rey[4], c, vx * vz);
    if (s.rey[5]) upd(s.rey[5], c, vy * vz);
    if (s.vort[0] || s.vort[1] || s.vort[2] || s.vort[3]) {
        const int x = int(c % uint32_t(k.nx)), y = int(c / uint32_t(k.nx) % uint32_t(k.ny));
        const int z = int(c / k.sxy);
        float wx = 0.f, wy = 0.f, wz = 0.f;
        if (is_interior(k, x, y, z)) {
            const uint32_t xp = c + 1, xm = c - 1, yp = c + uint32_t(k.nx), ym = c - uint32_t(k.nx);
            const uint32_t zp = c + k.sxy, zm = c - k.sxy;
            wx = 0.5f * (uz[yp] - uz[ym]) - 0.5f * (uy[zp] - uy[zm]);
            wy = 0.5f * (ux[zp] - ux[zm]) - 0.5f * (uz[xp] - uz[xm]);
            wz = 0.5f * (uy[xp] - uy[xm]) - 0.5f * (ux[yp] - ux[ym]);
        }
        if (s.vort[0]) s.vort[0][c] = wx;
        if (s.vort[1]) s.vort[1][c] = wy;
        if (s.vort[2]) s.vort[2][c] = wz;
        if (s.vort[3]) s.vort[3][c] = __builtin_sqrtf(wx * wx + wy * wy + wz * wz);
    }
}

// K1 for every jet.  A block takes 256 consecutive cells of an x-y plane
// and marches over kPhaseLZ planes (phase_column): jet j owns blocks
// [cell_block_begin_j, cell_block_begin_{j+1}), numbered plane block fastest,
// then z-chunk, ceil(nx ny / 256) ceil((nz-2) / kPhaseLZ) per jet.  Flat
// 256-cell rows keep every block's accesses contiguous for any box width;
// blocks run in launch order (an XCD-contiguous order measured 1.5x slower
// on the ensemble's narrow boxes).
__global__ void __launch_bounds__(256) k_phase_batch(const JetSlot *__restrict__ jets, int njets) {
    const uint32_t b = blockIdx.x;
    const JetSlot &js = jets[find_slot(jets, njets, b, 1)];
    const KParams &k = js.k;
    const uint32_t lb = b - js.cell_block_begin, pb = (k.sxy + 255u) / 256u;
    const uint32_t zc = lb / pb, cxy = (lb - zc * pb) * 256u + threadIdx.x;
    if (cxy >= k.sxy) return;
    const int x = int(cxy % uint32_t(k.nx)), y = int(cxy / uint32_t(k.nx));
    const int z0 = 1 + int(zc) * kPhaseLZ, z1 = min(z0 + kPhaseLZ, k.nz - 1);
    if (y < 1 || y > k.ny - 2) return;
    if (x == 0 || x == k.nx - 1) {
        for (int z = z0; z < z1; ++z) phase_keep(js.phi, cell_id(k, x, y, z));
        return;
    }
    phase_column(k, js.g, js.phi, x, y, z0, z1);
}

// K1 for every jet when every row length is a multiple of 4: k_phase_quad's
// quads, jet j owning blocks [cell_block_begin_j, cell_block_begin_{j+1}),
// numbered quad block fastest, then z-chunk of kEnsemblePhaseLZ planes,
// ceil((nx/4)(ny-2) / 256) ceil((nz-2) / kEnsemblePhaseLZ) per jet, in launch
// order.
constexpr int kEnsemblePhaseLZ = 4;
__global__ void __launch_bounds__(256) k_phase_quad_batch(const JetSlot *__restrict__ jets, int njets) {
    const uint32_t b = blockIdx.x;
    const JetSlot &js = jets[find_slot(jets, njets, b, 1)];
    const KParams &k = js.k;
    const uint32_t rowq = (uint32_t(k.nx) >> 2) * uint32_t(k.ny - 2), pb = (rowq + 255u) >> 8;
    const uint32_t lb = b - js.cell_block_begin, zc = lb / pb, q = (lb - zc * pb) * 256u + threadIdx.x;
    if (q >= rowq) return;
    // phi non-temporally where the members' g together exceed the Infinity
    // Cache (g_cached is set for the whole ensemble): +0.5% on the 32 jets
    if (k.g_cached)
        phase_quad_column<kEnsemblePhaseLZ, false>(k, js.g, js.phi, q, zc);
    else
        phase_quad_column<kEnsemblePhaseLZ, true>(k, js.g, js.phi, q, zc);
}
#endif  // SSMD_STEP_KERNELS


// Persistent lattice kernel: gridDim.x resident blocks take the tiles of all
// jets from a queue, in tile order (queue[0] = next ticket), so blocks that
// drew cheap tiles take more of them.  A block leaves when its ticket is past
// the last tile; the last block to leave (queue[1] counts them) resets both
// words, so every launch -- a graph replay included -- starts from 0.
// LDS of the 128 x 4 persistent kernel, which also runs 64 x 8 tiles (the
// same 512 threads) for members whose rows a 64-wide tile covers with fewer
// idle lanes; 64 x 8 needs less LDS than 128 x 4.
template <int TX, int TY>
union MixedSmem {
    LatticeSmem<TX, TY> main;
    LatticeSmem<64, 8> alt;
};
template <int TX, int TY, bool MACROS> constexpr bool mixed_tiles() { return TX == 128 && TY == 4 && !MACROS; }

template <int Q, class P, int TX, int TY, bool MACROS>
__global__ void __launch_bounds__(TX *TY) k_lattice_batch(const JetSlot *__restrict__ jets, int njets,
                                                          uint32_t total_tiles, uint32_t *__restrict__ queue) {
    constexpr bool kMixed = mixed_tiles<TX, TY, MACROS>();
    __shared__ std::conditional_t<kMixed, MixedSmem<TX, TY>, LatticeSmem<TX, TY>> sm;
    __shared__ uint32_t ticket;
    using PT = typename P::T;
    const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
    for (;;) {
        __syncthreads();  // the previous tile's LDS readers and ticket readers are done
        if (lead) ticket = __hip_atomic_fetch_add(&queue[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t t = ticket;
        if (t >= total_tiles) break;
        const int j = find_slot(jets, njets, t, 0);
        const JetSlot &js = jets[j];
        uint32_t lt = t - js.tile_begin;
        const int bx = int(lt % uint32_t(js.tiles_x));
        lt /= uint32_t(js.tiles_x);
        const int by = int(lt % uint32_t(js.tiles_y));
        const int bz = int(lt / uint32_t(js.tiles_y));
        const int zb = 1 + bz * js.lz;
        const int ze = min(zb + js.lz, js.k.nz - 1);
        if constexpr (kMixed) {
            if (js.narrow) {
                // the 512 threads as 64 x 8: a wavefront (64 consecutive x of one
                // threadIdx.y) stays one row
                const int l = int(threadIdx.x) + TX * int(threadIdx.y);
                lattice_tile<Q, P, 64, 8, MACROS>(js.k, static_cast<const PT *>(js.f), static_cast<PT *>(js.f_next),
                                                  js.g_next, js.phi, &js.mac, bx * 64, by * 8, zb, ze, sm.alt,
                                                  l & 63, l >> 6);
            } else {
                lattice_tile<Q, P, TX, TY, MACROS>(js.k, static_cast<const PT *>(js.f), static_cast<PT *>(js.f_next),
                                                   js.g_next, js.phi, &js.mac, bx * TX, by * TY, zb, ze, sm.main);
            }
        } else {
            lattice_tile<Q, P, TX, TY, MACROS>(js.k, static_cast<const PT *>(js.f), static_cast<PT *>(js.f_next),
                                               js.g_next, js.phi, &js.mac, bx * TX, by * TY, zb, ze, sm);
        }
    }
    if (lead &&
        __hip_atomic_fetch_add(&queue[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
        // every block has drawn its last ticket: reset for the next launch
        __hip_atomic_store(&queue[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&queue[1], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// --------------------------------------------------------------------------
// Boundary kernels.  Each writes the SOURCE slots of the populations the
// reference writes at their arrival cells, so arrival a along c_i maps to
// slot (a - c_i).

// Seeded inlet noise: SplitMix64 chained over (seed, step, cell, component).
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float noise_uniform(uint64_t seed, uint64_t step, uint64_t cell, uint64_t comp) {
    const uint64_t h = splitmix64(splitmix64(splitmix64(seed ^ step) ^ cell) ^ comp);
    return float(uint32_t(h >> 40)) * 1.1920928955078125e-7f - 1.0f;
}

// Non-equilibrium reconstruction at a fluid node (no force term,
// BoundaryConditions.cuh:87-94): feq(rho, u_bc) + omco * fneq(Pi, u at node).
template <int Q, int I>
__device__ __forceinline__ float bc_population(const KParams &k, const Macro &mac, uint32_t fl, float rho, float ux,
                                               float uy, float uz, float omco) {
    constexpr Dir c = kDir[I];
    constexpr float w = weight<Q>(I);
    const float uu = 1.5f * (ux * ux + uy * uy + uz * uz);
    const float cu = 3.0f * (float(c.x) * ux + float(c.y) * uy + float(c.z) * uz);
    float feq = rho * (1.0f - uu + cu + 0.5f * cu * cu);
    if constexpr (Q == 27) feq += rho * ((1.0f / 6.0f) * cu * cu * cu - uu * cu);
    feq = w * feq - w;
    const float pxx = mac.pxx[fl], pyy = mac.pyy[fl], pzz = mac.pzz[fl];
    const float pxy = mac.pxy[fl], pxz = mac.pxz[fl], pyz = mac.pyz[fl];
    float s = (float(c.x * c.x) - kCs2) * pxx + (float(c.y * c.y) - kCs2) * pyy + (float(c.z * c.z) - kCs2) * pzz +
              2.0f * (float(c.x * c.y) * pxy + float(c.x * c.z) * pxz + float(c.y * c.z) * pyz);
    if constexpr (Q == 27) {
        const float nux = mac.ux[fl], nuy = mac.uy[fl], nuz = mac.uz[fl];
        constexpr int kxxy = 3 * c.x * c.x * c.y - c.y;
        constexpr int kxxz = 3 * c.x * c.x * c.z - c.z;
        constexpr int kxyy = 3 * c.x * c.y * c.y - c.x;
        constexpr int kyyz = 3 * c.y * c.y * c.z - c.z;
        constexpr int kxzz = 3 * c.x * c.z * c.z - c.x;
        constexpr int kyzz = 3 * c.y * c.z * c.z - c.y;
        constexpr int kxyz = 6 * c.x * c.y * c.z;
        s += float(kxxy) * (pxx * nuy + 2.0f * nux * pxy) + float(kxxz) * (pxx * nuz + 2.0f * nux * pxz) +
             float(kxyy) * (pyy * nux + 2.0f * nuy * pxy) + float(kyyz) * (pyy * nuz + 2.0f * nuy * pyz) +
             float(kxzz) * (pzz * nux + 2.0f * nuz * pxz) + float(kyzz) * (pzz * nuy + 2.0f * nuz * pyz) +
             float(kxyz) * (nux * pyz + nuy * pxz + nuz * pxy);
    }
    return feq + omco * ((w * 4.5f) * s);
}

// Device step index for graph replays: counter if given, else the argument.
__device__ __forceinline__ uint64_t step_index(const uint64_t *counter, uint64_t step) {
    return counter ? *counter : step;
}
__device__ __forceinline__ void bump(uint64_t *counter, int doit) {
    if (doit && counter && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 &&
        threadIdx.y == 0)
        *counter += 1;
}

// The inlet and outflow kernels spread one face cell over several threads:
// layer `slot` (blockIdx.z) of the launch emits the slot-th population of
// the Q that cross the face (component `SIGN` along `AXIS`), and layer 0 also
// writes the cell's boundary moments and phase moments.
template <int Q, int AXIS, int SIGN>
constexpr int crossing_rank(int i) {
    int r = 0;
    for (int j = 0; j < i; ++j) r += (AXIS == 1 ? kDir[j].y : kDir[j].z) == SIGN;
    return r;
}
template <int Q>
constexpr int crossing_count() {
    return crossing_rank<Q, 2, 1>(Q);  // the same on every axis and side: Q19 5, Q27 9
}
static_assert(crossing_count<19>() == 5 && crossing_count<27>() == 9, "populations crossing a face");

// applyWaterInflow (axis 1, face y=0) / applyOilInflow (axis 2, face z=0).
template <int Q, class P, int AXIS>
__device__ __forceinline__ void inflow_face(const KParams &k, typename P::T *__restrict__ fout, float *__restrict__ gout,
                                            float *__restrict__ phi, const Macro &mac, const uint64_t *counter,
                                            uint64_t step, int a, int b, int slot) {
    const JetK &j = AXIS == 1 ? k.water : k.oil;
    const int nb = AXIS == 1 ? k.nz : k.ny;
    if (a < k.nx && b < nb) {
        const float da = float(a) - j.ca, db = float(b) - j.cb;
        if (da * da + db * db <= j.r2) {
            const int x = a, y = AXIS == 1 ? 0 : b, z = AXIS == 1 ? b : 0;
            const uint32_t bnd = cell_id(k, x, y, z);
            float u3[3] = {0.f, 0.f, 0.f};
            if (k.perturb != 0.f) {
                const uint64_t s = step_index(counter, step);
                const float amp = k.perturb * j.u;
                for (int q = 0; q < 3; ++q) u3[q] = amp * noise_uniform(k.seed, s, bnd, uint64_t(q));
            }
            u3[AXIS] += j.u;
            const float ux = u3[0], uy = u3[1], uz = u3[2];
            const float rho = 1.0f, ph = j.phi;
            if (slot == 0) {
                mac.rho[bnd] = rho;
                phi[bnd] = ph;
                mac.ux[bnd] = ux;
                mac.uy[bnd] = uy;
                mac.uz[bnd] = uz;
                // population +axis of the source bnd, w (ph (1 + 4 un)), in moment form
                const float un = AXIS == 1 ? uy : uz;
                gout[bnd] = ph * (1.0f + 4.0f * un);
                gout[size_t(AXIS + 1) * k.N + bnd] = 0.f;
            }
            const float omco = AXIS == 1 ? k.omco_water : k.omco_oil;
            sfor<0, Q>([&](auto I) {
                constexpr int i = I;
                constexpr Dir c = kDir[i];
                if constexpr ((AXIS == 1 ? c.y : c.z) == 1) {
                    if (slot != crossing_rank<Q, AXIS, 1>(i)) return;
                    const int xx = x + c.x, yy = y + c.y, zz = z + c.z;
                    bool ok;
                    if constexpr (AXIS == 2)
                        ok = xx >= 1 && xx <= k.nx - 2 && yy >= 1 && yy <= k.ny - 2;
                    else
                        ok = xx >= 0 && xx < k.nx && zz >= 0 && zz < k.nz;
                    if (ok) {
                        const uint32_t fl = cell_id(k, xx, yy, zz);
                        fout[size_t(i) * k.N + bnd] = P::store(bc_population<Q, i>(k, mac, fl, rho, ux, uy, uz, omco));
                    }
                }
            });
        }
    }
}

// applyOutflowY (axis 1) / applyOutflowZ (axis 2).
template <int Q, class P, int AXIS>
__device__ __forceinline__ void outflow_face(const KParams &k, typename P::T *__restrict__ fout,
                                             float *__restrict__ gout, float *__restrict__ phi, const Macro &mac,
                                             int a, int b, int slot) {
    const int nb = AXIS == 1 ? k.nz : k.ny;
    if (a >= 1 && a <= k.nx - 2 && b >= 1 && b <= nb - 2) {
        const int x = a, y = AXIS == 1 ? k.ny - 1 : b, z = AXIS == 1 ? b : k.nz - 1;
        const uint32_t bnd = cell_id(k, x, y, z);
        const uint32_t in = AXIS == 1 ? cell_id(k, x, k.ny - 2, z) : cell_id(k, x, y, k.nz - 2);
        const float rho = mac.rho[in], ph = phi[in], ux = mac.ux[in], uy = mac.uy[in], uz = mac.uz[in];
        if (slot == 0) {
            mac.rho[bnd] = rho;
            phi[bnd] = ph;
            mac.ux[bnd] = ux;
            mac.uy[bnd] = uy;
            mac.uz[bnd] = uz;
            // population -axis of the source bnd, w (ph (1 - 4 un)), in moment form
            const float un = AXIS == 1 ? uy : uz;
            gout[bnd] = ph * (1.0f - 4.0f * un);
            gout[size_t(AXIS + 1) * k.N + bnd] = 0.f;
        }
        sfor<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr Dir c = kDir[i];
            if constexpr ((AXIS == 1 ? c.y : c.z) == -1) {
                if (slot != crossing_rank<Q, AXIS, -1>(i)) return;
                const int xx = x + c.x;
                const int yy = AXIS == 1 ? k.ny - 2 : y + c.y;
                const int zz = AXIS == 1 ? z + c.z : k.nz - 2;
                bool ok = true;
                if constexpr (AXIS == 1) ok = xx >= 1 && xx <= k.nx - 2 && zz >= 1 && zz <= k.nz - 2;
                if (ok) {
                    const uint32_t fl = cell_id(k, xx, yy, zz);
                    fout[size_t(i) * k.N + bnd] = P::store(bc_population<Q, i>(k, mac, fl, rho, ux, uy, uz, k.omco_ref));
                }
            }
        });
    }
}

// periodicX (BoundaryConditions.cuh:295-349) and its mirrors on y and z.
// Face coordinates (a, b) run over the two other axes in increasing order.
template <int AXIS>
__device__ __forceinline__ void periodic_face(const KParams &k, float *__restrict__ phi, const Macro &mac, int a, int b,
                                              bool with_u) {
    // periodicX's ghost wrap of phi and u (BoundaryConditions.cuh:334-348) and
    // its mirrors.  The population copies of the same kernel are folded into
    // the pull addresses of k_phase / k_lattice (wrap_shift), so the slots
    // they would write are never read and are not written here.
    constexpr int A1 = AXIS == 0 ? 1 : 0, A2 = AXIS == 2 ? 1 : 2;
    const int n[3] = {k.nx, k.ny, k.nz};
    if (a < n[A1] && b < n[A2]) {
        const int nn = n[AXIS];
        auto at = [&](int s, int aa, int bb) {
            int r[3];
            r[AXIS] = s;
            r[A1] = aa;
            r[A2] = bb;
            return cell_id(k, r[0], r[1], r[2]);
        };
        const uint32_t bL = at(1, a, b), bR = at(nn - 2, a, b), gL = at(0, a, b), gR = at(nn - 1, a, b);
        phi[gL] = phi[bR];
        phi[gR] = phi[bL];
        if (with_u) {  // u on the ghost layer is read by no kernel; kept current on macro steps
            mac.ux[gL] = mac.ux[bR];
            mac.ux[gR] = mac.ux[bL];
            mac.uy[gL] = mac.uy[bR];
            mac.uy[gR] = mac.uy[bL];
            mac.uz[gL] = mac.uz[bR];
            mac.uz[gR] = mac.uz[bL];
        }
    }
}

// periodicX then periodicY (the y mirror), as one pass over the ghost ring
// of the x and y faces.  Position a < ny is the x-ghost pair of row y = a
// (all rows, ghost rows included); a >= ny is the y-ghost pair of column
// x = a - ny + 1 (interior columns).  Run in that order, the two copies leave
// every ghost cell of the ring equal to one cell of the interior columns and
// rows: its own row/column wrapped on the ghost axis -- for the four corner
// columns wrapped on both axes, since periodicY copies the row periodicX has
// just filled.  Every source is read here and written by no thread of this
// pass, so all cells are written in one pass with the sequential result.
// `sb` is the plane the sources are read from: b, or nz-2 for b = nz-1 when
// the z outflow's copy of plane nz-2 into nz-1 runs in the same launch
// (k_faces_z); the sources are interior (x, y), which that copy covers.
__device__ __forceinline__ void periodic_xy(const KParams &k, float *__restrict__ phi, const Macro &mac, int a, int b,
                                            bool with_u, int sb) {
    if (a >= k.ny + k.nx - 2 || b >= k.nz) return;
    uint32_t gL, gR, sL, sR;
    if (a < k.ny) {
        const int ys = a == 0 ? k.ny - 2 : (a == k.ny - 1 ? 1 : a);  // the row periodicY leaves
        gL = cell_id(k, 0, a, b);
        gR = cell_id(k, k.nx - 1, a, b);
        sL = cell_id(k, k.nx - 2, ys, sb);
        sR = cell_id(k, 1, ys, sb);
    } else {
        const int x = a - k.ny + 1;
        gL = cell_id(k, x, 0, b);
        gR = cell_id(k, x, k.ny - 1, b);
        sL = cell_id(k, x, k.ny - 2, sb);
        sR = cell_id(k, x, 1, sb);
    }
    phi[gL] = phi[sL];
    phi[gR] = phi[sR];
    if (with_u) {  // u on the ghost layer is read by no kernel; kept current on macro steps
        mac.ux[gL] = mac.ux[sL];
        mac.ux[gR] = mac.ux[sR];
        mac.uy[gL] = mac.uy[sL];
        mac.uy[gR] = mac.uy[sR];
        mac.uz[gL] = mac.uz[sL];
        mac.uz[gR] = mac.uz[sR];
    }
}

// Face operations, in the order a step applies them (main.cu:165-169, then
// the new y and z periodic copies).
enum FaceKind {
    kInflowY = 0,   // applyWaterInflow
    kInflowZ,       // applyOilInflow
    kOutflowY,      // applyOutflowY
    kOutflowZ,      // applyOutflowZ
    kPeriodicXRef,  // periodicX
    kPeriodicXTrue,
    kPeriodicYRef,
    kPeriodicYTrue,
    kPeriodicZRef,
    kPeriodicZTrue,
    kPeriodicXY,    // periodicX followed by periodicY, in one launch (periodic_xy)
};

// Block layers (blockIdx.z) of a face kind's launch: one per crossing
// population for inlets and outflows, one otherwise.
template <int Q>
constexpr int face_layers(int kind) {
    return kind == kInflowY || kind == kInflowZ || kind == kOutflowY || kind == kOutflowZ ? crossing_count<Q>() : 1;
}

// The face (block range) a kind runs over: 0, 1, 2 = normal to x, y, z;
// 3 = the ghost ring of the x and y faces (periodic_xy).
constexpr int face_axis(int kind) {
    return kind == kPeriodicXY ? 3
           : kind == kInflowY || kind == kOutflowY || kind == kPeriodicYRef || kind == kPeriodicYTrue
               ? 1
               : (kind == kPeriodicXRef || kind == kPeriodicXTrue ? 0 : 2);
}

// Extent (na, nb) of a face normal to `axis`: the two other axes in order.
__device__ __host__ __forceinline__ void face_extent(int axis, int nx, int ny, int nz, int &na, int &nb) {
    if (axis == 3) {  // ring positions (periodic_xy) x planes
        na = ny + nx - 2;
        nb = nz;
        return;
    }
    na = axis == 0 ? ny : nx;
    nb = axis == 2 ? ny : nz;
}

template <int Q, class P, int KIND>
__device__ __forceinline__ void face_op(const KParams &k, typename P::T *fout, float *gout, float *phi,
                                        const Macro &mac, const uint64_t *counter, uint64_t step, int a, int b,
                                        bool macros, int slot) {
    if constexpr (KIND == kInflowY) inflow_face<Q, P, 1>(k, fout, gout, phi, mac, counter, step, a, b, slot);
    if constexpr (KIND == kInflowZ) inflow_face<Q, P, 2>(k, fout, gout, phi, mac, counter, step, a, b, slot);
    if constexpr (KIND == kOutflowY) outflow_face<Q, P, 1>(k, fout, gout, phi, mac, a, b, slot);
    if constexpr (KIND == kOutflowZ) outflow_face<Q, P, 2>(k, fout, gout, phi, mac, a, b, slot);
    if constexpr (KIND == kPeriodicXRef) periodic_face<0>(k, phi, mac, a, b, macros);
    if constexpr (KIND == kPeriodicXTrue) periodic_face<0>(k, phi, mac, a, b, macros);
    if constexpr (KIND == kPeriodicYRef) periodic_face<1>(k, phi, mac, a, b, macros);
    if constexpr (KIND == kPeriodicYTrue) periodic_face<1>(k, phi, mac, a, b, macros);
    if constexpr (KIND == kPeriodicZRef) periodic_face<2>(k, phi, mac, a, b, macros);
    if constexpr (KIND == kPeriodicZTrue) periodic_face<2>(k, phi, mac, a, b, macros);
    if constexpr (KIND == kPeriodicXY) periodic_xy(k, phi, mac, a, b, macros, b);
}

// flags: bit 0 = advance the device step counter (last face kernel of the
// step), bit 1 = macro step (keep u current on periodic ghost layers).
// One field set: 2D grid of 64x4 blocks over the face.  The inlet kinds read
// the device step counter, so only the other kinds may advance it.
template <int Q, class P, int KIND>
__global__ void __launch_bounds__(256) k_face(KParams k, typename P::T *fout, float *gout, float *phi, Macro mac,
                                              uint64_t *counter, uint64_t step, int flags) {
    const int a = blockIdx.x * 64 + threadIdx.x;
    const int b = blockIdx.y * 4 + threadIdx.y;
    face_op<Q, P, KIND>(k, fout, gout, phi, mac, counter, step, a, b, (flags & 2) != 0, int(blockIdx.z));
    if constexpr (KIND != kInflowY && KIND != kInflowZ) bump(counter, flags & 1);
}

// Ensemble: 1D grid of 64x4 blocks over every jet's face blocks.
template <int Q, class P, int KIND>
__global__ void __launch_bounds__(256) k_face_batch(const JetSlot *__restrict__ jets, int njets, uint64_t *counter,
                                                    uint64_t step, int flags) {
    constexpr int axis = face_axis(KIND);
    const int j = find_slot(jets, njets, blockIdx.x, 2 + axis);
    const JetSlot &js = jets[j];
    int na, nb;
    face_extent(axis, js.k.nx, js.k.ny, js.k.nz, na, nb);
    const int bxs = (na + 63) / 64;
    const uint32_t lb = blockIdx.x - js.face_begin[axis];
    const int a = int(lb % uint32_t(bxs)) * 64 + threadIdx.x;
    const int b = int(lb / uint32_t(bxs)) * 4 + threadIdx.y;
    using PT = typename P::T;
    face_op<Q, P, KIND>(js.k, static_cast<PT *>(js.f_next), js.g_next, js.phi, js.mac, counter, step, a, b,
                        (flags & 2) != 0, int(blockIdx.z));
    if constexpr (KIND != kInflowY && KIND != kInflowZ) bump(counter, flags & 1);
}

// The face operations of a step with faces among {z inlet, z outflow, x-y
// ghost ring} in ONE launch (graph steps, which write no ghost u), each over
// its own range of 64x4 blocks: `nin` blocks of the inlet (face blocks times
// its crossing-population layers), `nout` of the outflow, then the ring.
// Run one after another, the ring copies plane nz-1 after the outflow has
// copied plane nz-2 into it; here the ring reads plane nz-2 for those cells
// (periodic_xy's `sb`), the values that copy writes.  The inlet disc keeps
// clear of the x and y faces (ssmd_validate), so the inlet and the ring touch
// disjoint cells, as do the inlet (planes 0, 1) and the outflow (nz-2, nz-1).
// The inlet reads the device step counter; the phase kernel of the step has
// advanced it.  Results are those of the kernels launched in order.
template <int Q, class P, bool IN, bool OUT, bool RING>
__global__ void __launch_bounds__(256) k_faces_z(KParams k, typename P::T *fout, float *gout, float *phi, Macro mac,
                                                 const uint64_t *counter, uint32_t nin, uint32_t nout) {
    uint32_t b = blockIdx.x;
    const uint32_t fx = uint32_t(k.nx + 63) / 64, fxy = fx * (uint32_t(k.ny + 3) / 4);  // z-face blocks per layer
    if constexpr (IN) {
        if (b < nin) {
            const uint32_t r = b % fxy;
            inflow_face<Q, P, 2>(k, fout, gout, phi, mac, counter, 0, int(r % fx) * 64 + int(threadIdx.x),
                                 int(r / fx) * 4 + int(threadIdx.y), int(b / fxy));
            return;
        }
        b -= nin;
    }
    if constexpr (OUT) {
        if (b < nout) {
            const uint32_t r = b % fxy;
            outflow_face<Q, P, 2>(k, fout, gout, phi, mac, int(r % fx) * 64 + int(threadIdx.x),
                                  int(r / fx) * 4 + int(threadIdx.y), int(b / fxy));
            return;
        }
        b -= nout;
    }
    if constexpr (RING) {
        const uint32_t rx = uint32_t(k.ny + k.nx - 2 + 63) / 64;
        const int a = int(b % rx) * 64 + int(threadIdx.x), bb = int(b / rx) * 4 + int(threadIdx.y);
        periodic_xy(k, phi, mac, a, bb, false, OUT && bb == k.nz - 1 ? k.nz - 2 : bb);
    }
}

#ifdef SSMD_STEP_KERNELS
__global__ void k_set_counter(uint64_t *counter, uint64_t v) { *counter = v; }
// Advances the device step index when no outflow/periodic kernel follows
// the inlets (the inlet kernels read the counter, so they never write it).
__global__ void k_bump(uint64_t *counter) { *counter += 1; }
#endif  // SSMD_STEP_KERNELS


// --------------------------------------------------------------------------
// t = 0 state (initialConditions.cu:33-140) in the source-indexed layout:
// slot [i][s] = reference f[i][s + c_i] = feq_i(rho, u at s + c_i).
__device__ __forceinline__ void init_macros(const KParams &k, int x, int y, int z, float &phi, float &ux, float &uy,
                                            float &uz) {
    phi = 0.f;
    ux = uy = uz = 0.f;
    if (k.water.enabled && y == 0) {
        const float da = float(x) - k.water.ca, db = float(z) - k.water.cb;
        if (da * da + db * db <= k.water.r2) {
            phi = k.water.phi;
            uy = k.water.u;
        }
    }
    if (k.oil.enabled && z == 0) {
        const float da = float(x) - k.oil.ca, db = float(y) - k.oil.cb;
        if (da * da + db * db <= k.oil.r2) {
            phi = k.oil.phi;
            uz = k.oil.u;
        }
    }
}

template <int Q, class P>
__global__ void __launch_bounds__(256) k_init(KParams k, typename P::T *f0, typename P::T *f1, float *g0, float *g1,
                                              float *phi, Macro mac) {
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= k.N) return;
    const int x = int(c % uint32_t(k.nx));
    const uint32_t r = c / uint32_t(k.nx);
    const int y = int(r % uint32_t(k.ny));
    const int z = int(r / uint32_t(k.ny));
    float ph, ux, uy, uz;
    init_macros(k, x, y, z, ph, ux, uy, uz);
    phi[c] = ph;
    mac.rho[c] = 1.0f;
    mac.ux[c] = ux;
    mac.uy[c] = uy;
    mac.uz[c] = uz;
    mac.pxx[c] = mac.pyy[c] = mac.pzz[c] = mac.pxy[c] = mac.pxz[c] = mac.pyz[c] = 0.f;
    mac.fsx[c] = mac.fsy[c] = mac.fsz[c] = 0.f;
    sfor<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr Dir d = kDir[i];
        constexpr float w = weight<Q>(i);
        const int ax = x + d.x, ay = y + d.y, az = z + d.z;
        float v = 0.f;
        if (ax >= 0 && ax < k.nx && ay >= 0 && ay < k.ny && az >= 0 && az < k.nz) {
            float aph, aux, auy, auz;
            init_macros(k, ax, ay, az, aph, aux, auy, auz);
            const float uu = 1.5f * (aux * aux + auy * auy + auz * auz);
            const float cu = 3.0f * (float(d.x) * aux + float(d.y) * auy + float(d.z) * auz);
            float e = 1.0f - uu + cu + 0.5f * cu * cu;
            if constexpr (Q == 27) e += (1.0f / 6.0f) * cu * cu * cu - uu * cu;
            v = w * e - w;
        }
        const typename P::T sv = P::store(v);
        f0[size_t(i) * k.N + c] = sv;
        f1[size_t(i) * k.N + c] = sv;
    });
    // Phase populations in moment form: all zero.  Every population an
    // interior cell pulls at the first step is the t = 0 equilibrium of that
    // interior cell (initialConditions.cu sets g to the local equilibrium),
    // and interior cells start at phi = 0, u = 0, so each one is zero.
    // Slots no interior cell pulls are never read.
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        g0[size_t(kk) * k.N + c] = 0.f;
        g1[size_t(kk) * k.N + c] = 0.f;
    }
}

// --------------------------------------------------------------------------
// FP64 reductions over interior cells: sum phi, rho*ux, rho*uy, rho*uz.
// Pass 1 writes one partial per block; pass 2 (one block) sums them in a
// fixed order, so the result is deterministic.
#ifdef SSMD_STEP_KERNELS
__global__ void __launch_bounds__(256) k_reduce1(KParams k, const float *phi, Macro mac, double *part) {
    __shared__ double sh[4][256];
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < k.N; c += gridDim.x * 256u) {
        const int x = int(c % uint32_t(k.nx));
        const uint32_t r = c / uint32_t(k.nx);
        const int y = int(r % uint32_t(k.ny));
        const int z = int(r / uint32_t(k.ny));
        if (!is_interior(k, x, y, z)) continue;
        const double rho = mac.rho[c];
        s0 += double(phi[c]);
        s1 += rho * double(mac.ux[c]);
        s2 += rho * double(mac.uy[c]);
        s3 += rho * double(mac.uz[c]);
    }
    sh[0][threadIdx.x] = s0;
    sh[1][threadIdx.x] = s1;
    sh[2][threadIdx.x] = s2;
    sh[3][threadIdx.x] = s3;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_reduce2(const double *part, int nparts, double *out) {
    __shared__ double sh[4][256];
    double s[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += 256)
        for (int q = 0; q < 4; ++q) s[q] += part[b * 4 + q];
    for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = sh[threadIdx.x][0];
}

// Source coordinate of arrival coordinate v along one axis for a direction
// component c, with the periodic image folded in as the pulls fold it
// (wrap_shift); -1 where the source lies outside the box.
__device__ __forceinline__ int export_source(int v, int c, int n, int wrap) {
    int s = v - c;
    if (s < 0 || s >= n) return -1;
    if (wrap && s == 0 && c > 0) s = wrap_shift(wrap, n);
    if (wrap && s == n - 1 && c < 0) s = n - 1 - wrap_shift(wrap, n);
    return s;
}

// The populations in the reference's arrival-indexed layout (LBMFields f
// and g, LBMFields.cuh:26-85): fo[i][a] = f[i][a - c_i] in the storage
// format, go[k][a] = w_k (g_phi + c_k . o) of source a - c_k formed with the
// phase kernel's FP32 operations (phase_sum), 0 where an arrival has no
// source in the box.  One thread per arrival cell.
template <int Q, class P>
__global__ void __launch_bounds__(256) k_export(KParams k, const typename P::T *__restrict__ f,
                                                const float *__restrict__ g, typename P::T *__restrict__ fo,
                                                float *__restrict__ go) {
    using PT = typename P::T;
    const uint32_t a = blockIdx.x * 256u + threadIdx.x;
    if (a >= k.N) return;
    const int x = int(a % uint32_t(k.nx)), y = int(a / uint32_t(k.nx) % uint32_t(k.ny)), z = int(a / k.sxy);
    const size_t N = k.N;
    auto source = [&](int cx, int cy, int cz) -> long {
        const int sx = export_source(x, cx, k.nx, k.wrap_x), sy = export_source(y, cy, k.ny, k.wrap_y),
                  sz = export_source(z, cz, k.nz, k.wrap_z);
        return sx < 0 || sy < 0 || sz < 0 ? -1 : long(cell_id(k, sx, sy, sz));
    };
    if (fo) {
        sfor<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr Dir d = kDir[i];
            const long s = source(d.x, d.y, d.z);
            fo[size_t(i) * N + a] = s < 0 ? PT(0) : f[size_t(i) * N + size_t(s)];
        });
    }
    if (go) {
        go[a] = weight7(0) * g[a];
        const float *o[3] = {g + N, g + 2 * N, g + 3 * N};
        for (int ax = 0; ax < 3; ++ax)
            for (int sg = 0; sg < 2; ++sg) {
                const int c = sg == 0 ? 1 : -1;
                const long s = source(ax == 0 ? c : 0, ax == 1 ? c : 0, ax == 2 ? c : 0);
                const int kk = 1 + 2 * ax + sg;  // +x, -x, +y, -y, +z, -z
                float v = 0.f;
                if (s >= 0) v = weight7(kk) * (sg == 0 ? g[s] + o[ax][s] : g[s] - o[ax][s]);
                go[size_t(kk) * N + a] = v;
            }
    }
}
#endif  // SSMD_STEP_KERNELS

template <class P>
__global__ void k_quantize(const float *in, float *out, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = P::load(P::store(in[i]));
}

}  // namespace ssmd
