// ssmd_oracle.cpp -- CPU restatement of the reference per-timestep hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is
// compared with; it is never linked into, loaded by or called from the
// product library.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it.
//
// It restates the arithmetic of brenogemelgo/ssmdLBM in the reference's
// own kernel order and its own (arrival-indexed, push) population layout,
// with the arithmetic type T selectable (double = the FP64 oracle, float =
// the reference's scalar_t and the CPU baseline):
//   init    lbm::setWaterJet / setOilJet / setInitialDensity / setDistros
//           src/initialConditions.cu:33-140 (launch order main.cu:109-112)
//   K1      phase::computePhase     src/phaseField.cuh:31-52
//   K2      phase::computeForces    src/phaseField.cuh:54-294
//   K3      lbm::computeMoments     src/lbm.cu:33-128
//   K4      lbm::streamCollide      src/lbm.cu:130-270
//   BCs     applyWaterInflow / applyOilInflow / applyOutflowY /
//           applyOutflowZ / periodicX  src/BoundaryConditions.cuh:36-349,
//           launched in that order after the graph (main.cu:162-169).
//   New (no reference function): periodicY (mirror of periodicX), the
//   TRUE-periodic wrap, and the seeded inlet-velocity noise.
// Velocity sets: src/velocitySet/D3Q19.cuh, D3Q27.cuh, D3Q7.cuh.
//
// Parity pinning: the functions here are checked by tests/test_oracle.py
// against closed forms and invariants (lattice moments, equilibrium and
// regularization identities, Guo-force moments, gradient exactness, sphere
// curvature, rest-state fixed point, shear-wave decay, conservation in a
// truly periodic box, mirror symmetries) and against the velocity-set
// tables extracted from the reference text (tests/golden/).
#include "ssmd.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <xmmintrin.h>

namespace {

// ---------------------------------------------------------------------------
// Velocity sets (direction order of D3Q19.cuh:82-131, D3Q27.cuh:92-140,
// D3Q7.cuh:74-122; weights D3Q19.cuh:50-80, D3Q27.cuh:50-90, D3Q7.cuh:50-72).
const int C27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},   {-1, 0, -1}, {0, 1, 1},   {0, -1, -1}, {1, -1, 0},
    {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},  {0, -1, 1},  {1, 1, 1},   {-1, -1, -1},
    {1, 1, -1},  {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};

double w_hydro(int q, int i) {
    if (q == 19) {
        if (i == 0) return 1.0 / 3.0;
        if (i <= 6) return 1.0 / 18.0;
        return 1.0 / 36.0;
    }
    if (i == 0) return 8.0 / 27.0;
    if (i <= 6) return 2.0 / 27.0;
    if (i <= 18) return 1.0 / 54.0;
    return 1.0 / 216.0;
}
double w_phase(int i) { return i == 0 ? 0.25 : 0.125; }

// ---------------------------------------------------------------------------
// Storage codecs for populations (cuda/utils.cuh:67-99 uses __float2half,
// round-to-nearest-even; bf16 is new).  Written from the IEEE-754 binary16
// and bfloat16 definitions.
uint16_t f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) return uint16_t(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
    if (ax >= 0x477FF000u) return uint16_t(sign | 0x7C00u);  // rounds past 65504
    if (ax >= 0x38800000u) {                                   // normal half
        const uint32_t base = ax - 0x38000000u;                // rebias exponent 127 -> 15
        uint32_t h = base >> 13;
        const uint32_t rem = base & 0x1FFFu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
        return uint16_t(sign | h);
    }
    const uint32_t e = ax >> 23;  // subnormal half: k * 2^-24
    if (e < 102u) return uint16_t(sign);
    const uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
    const uint32_t s = 126u - e;  // 14..24
    uint32_t k = m >> s;
    const uint32_t rem = m & ((1u << s) - 1u);
    const uint32_t half = 1u << (s - 1u);
    if (rem > half || (rem == half && (k & 1u))) ++k;
    return uint16_t(sign | k);
}

float f16_to_f32(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu;
    const uint32_t m = h & 0x3FFu;
    uint32_t x;
    if (e == 0x1Fu) {
        x = sign | 0x7F800000u | (m << 13);
    } else if (e != 0) {
        x = sign | ((e + 112u) << 23) | (m << 13);
    } else if (m == 0) {
        x = sign;
    } else {
        float v = float(m) * 5.9604644775390625e-8f;  // m * 2^-24, exact
        std::memcpy(&x, &v, 4);
        x |= sign;
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

uint16_t f32_to_bf16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7FFFFFFFu) > 0x7F800000u) return uint16_t((x >> 16) | 0x40u);
    x += 0x7FFFu + ((x >> 16) & 1u);
    return uint16_t(x >> 16);
}

float bf16_to_f32(uint16_t h) {
    uint32_t x = uint32_t(h) << 16;
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

template <class T>
T quantize(int fmt, T v) {
    if (fmt == SSMD_POP_F16) return T(f16_to_f32(f32_to_f16(float(v))));
    if (fmt == SSMD_POP_BF16) return T(bf16_to_f32(f32_to_bf16(float(v))));
    return v;
}

// ---------------------------------------------------------------------------
// Seeded inlet noise (new; no reference function): SplitMix64 finalizer
// (Steele, Lea, Flood 2014) chained over (seed, step, cell, component); the
// top 24 bits map exactly onto the float grid k*2^-23 - 1 in [-1, 1).
uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float noise_uniform(uint64_t seed, uint64_t step, uint64_t cell, uint64_t comp) {
    const uint64_t h = splitmix64(splitmix64(splitmix64(seed ^ step) ^ cell) ^ comp);
    return float(uint32_t(h >> 40)) * 1.1920928955078125e-7f - 1.0f;  // k * 2^-23 - 1
}

// ---------------------------------------------------------------------------
// Lattice functions, restated in the arithmetic type T.
template <class T>
struct Lat {
    int q;
    T w[27];
    T w7[7];
    T cs2, as2, oos;
    explicit Lat(int q_) : q(q_) {
        for (int i = 0; i < q; ++i) w[i] = T(w_hydro(q, i));
        for (int i = 0; i < 7; ++i) w7[i] = T(w_phase(i));
        cs2 = T(1.0 / 3.0);
        as2 = T(3);
        oos = T(1.0 / 6.0);
    }
    // D3Q19.cuh:133-140 / D3Q27.cuh:147-154
    T feq(int i, T rho, T uu, T cu) const {
        if (q == 19) return w[i] * rho * (T(1) - uu + cu + T(0.5) * cu * cu) - w[i];
        return w[i] * rho * (T(1) - uu + cu + T(0.5) * cu * cu + oos * cu * cu * cu - uu * cu) - w[i];
    }
    // D3Q19.cuh:142-161 / D3Q27.cuh:156-185
    T fneq(int i, T pxx, T pyy, T pzz, T pxy, T pxz, T pyz, T ux, T uy, T uz) const {
        const int cx = C27[i][0], cy = C27[i][1], cz = C27[i][2];
        T s = (T(cx * cx) - cs2) * pxx + (T(cy * cy) - cs2) * pyy + (T(cz * cz) - cs2) * pzz +
              T(2) * (T(cx * cy) * pxy + T(cx * cz) * pxz + T(cy * cz) * pyz);
        if (q == 27) {
            s += T(cx * cx * cx - cx) * (as2 * ux * pxx) + T(cy * cy * cy - cy) * (as2 * uy * pyy) +
                 T(cz * cz * cz - cz) * (as2 * uz * pzz) +
                 as2 * ((T(cx * cx * cy) - cs2 * T(cy)) * (pxx * uy + T(2) * ux * pxy) +
                        (T(cx * cx * cz) - cs2 * T(cz)) * (pxx * uz + T(2) * ux * pxz) +
                        (T(cx * cy * cy) - cs2 * T(cx)) * (pyy * ux + T(2) * uy * pxy) +
                        (T(cy * cy * cz) - cs2 * T(cz)) * (pyy * uz + T(2) * uy * pyz) +
                        (T(cx * cz * cz) - cs2 * T(cx)) * (pzz * ux + T(2) * uz * pxz) +
                        (T(cy * cz * cz) - cs2 * T(cy)) * (pzz * uy + T(2) * uz * pyz)) +
                 T(6) * T(cx * cy * cz) * (ux * pyz + uy * pxz + uz * pxy);
        }
        return (w[i] * T(4.5)) * s;
    }
    // D3Q19.cuh:163-177 (Guo forcing)
    T force(int i, T cu, T ux, T uy, T uz, T fx, T fy, T fz) const {
        const T cx = T(C27[i][0]), cy = T(C27[i][1]), cz = T(C27[i][2]);
        return w[i] * ((as2 * (cx - ux) + as2 * cu * cx) * fx + (as2 * (cy - uy) + as2 * cu * cy) * fy +
                       (as2 * (cz - uz) + as2 * cu * cz) * fz);
    }
    // D3Q7.cuh:124-132
    T geq(int k, T phi, T ux, T uy, T uz) const {
        return w7[k] * phi * (T(1) + T(4) * (T(C27[k][0]) * ux + T(C27[k][1]) * uy + T(C27[k][2]) * uz));
    }
    // D3Q7.cuh:140-148
    T antidiff(int k, T sharp, T nx, T ny, T nz) const {
        return w7[k] * sharp * (T(C27[k][0]) * nx + T(C27[k][1]) * ny + T(C27[k][2]) * nz);
    }
};

// Smagorinsky relaxation time (Hou et al. 1996): with nu_t = Cs^2 |S|,
// |S| = sqrt(2 S:S) and S = -Pi / (2 rho cs^2 tau) (cs^2 = 1/3), the total
// tau = tau0 + 3 nu_t solves tau^2 - tau0 tau - 4.5 Cs^2 sqrt(2 Pi:Pi) / rho = 0.
template <class T>
T smag_tau(T tau0, T pxx, T pyy, T pzz, T pxy, T pxz, T pyz, T rho, T cs) {
    const T q = pxx * pxx + pyy * pyy + pzz * pzz + T(2) * (pxy * pxy + pxz * pxz + pyz * pyz);
    return T(0.5) * (tau0 + std::sqrt(tau0 * tau0 + T(18) * cs * cs * std::sqrt(T(2) * q) / rho));
}

// ---------------------------------------------------------------------------
struct Base {
    virtual ~Base() = default;
    virtual void init() = 0;
    virtual void k_phase() = 0;
    virtual void k_forces() = 0;
    virtual void k_moments() = 0;
    virtual void k_collide() = 0;
    virtual void k_boundaries(uint64_t step) = 0;
    virtual double *field(int id) = 0;  // not used for float; see copy functions
    virtual int get_field(int id, double *out) = 0;
    virtual int set_field(int id, const double *in) = 0;
    virtual int get_pops(int which, double *out) = 0;
    virtual int set_pops(int which, const double *in) = 0;
    virtual void sums(double *out4) = 0;
    virtual void set_threads(int n) = 0;
    void step(uint64_t s) {
        k_phase();
        k_forces();
        k_moments();
        k_collide();
        k_boundaries(s);
    }
};

template <class T>
struct Oracle final : Base {
    ssmd_params p;
    Lat<T> L;
    int q;
    uint32_t nx, ny, nz;
    size_t N;
    int threads = 0;
    std::vector<T> fld[23];  // rho ux uy uz pxx pyy pzz pxy pxz pyz phi fsx fsy fsz,
                             // tau_t, feat0..feat7 (closure; no reference counterpart)
    std::vector<T> f, g;
    std::vector<T> nnx, nny, nnz, nind;  // K2 normals (tile storage in the reference)
    T visc_water, visc_oil, visc_diff, omco_water, omco_oil, omco_ref, sigma, gamma;
    T flux3;  // LEARNED flux closure: flux_coeff / 3, formed in double

    enum { RHO, UX, UY, UZ, PXX, PYY, PZZ, PXY, PXZ, PYZ, PHI, FSX, FSY, FSZ, TAU_T, FEAT0, NFIELDS = FEAT0 + 8 };

    explicit Oracle(const ssmd_params &pp) : p(pp), L(pp.q), q(pp.q), nx(pp.nx), ny(pp.ny), nz(pp.nz) {
        N = size_t(nx) * ny * nz;
        for (auto &v : fld) v.assign(N, T(0));
        f.assign(size_t(q) * N, T(0));
        g.assign(7 * N, T(0));
        nnx.assign(N, T(0));
        nny.assign(N, T(0));
        nnz.assign(N, T(0));
        nind.assign(N, T(0));
        // globalFunctions.cuh:232-283: viscosities and omega are formed in
        // double and cast to the arithmetic type.
        visc_water = T(p.visc_water);
        visc_oil = T(p.visc_oil);
        visc_diff = visc_oil - visc_water;
        auto omega = [](double nu) { return 1.0 / (0.5 + 3.0 * nu); };
        omco_water = T(1) - T(omega(p.visc_water));
        omco_oil = T(1) - T(omega(p.visc_oil));
        omco_ref = omco_water;
        sigma = T(p.sigma);
        gamma = T(p.gamma);
        flux3 = T(p.flux_coeff / 3.0);
    }

    void set_threads(int n) override {
        threads = n;
        flush_denormals();
    }
    // FP32 mode flushes denormals to zero (x86 MXCSR FTZ|DAZ) in every worker:
    // far-field populations decay into the denormal range, where x86 is slow.
    void flush_denormals() const {
        if (sizeof(T) != 4) return;
        _mm_setcsr(_mm_getcsr() | 0x8040u);
#pragma omp parallel num_threads(nthreads())
        { _mm_setcsr(_mm_getcsr() | 0x8040u); }
    }
    int nthreads() const {
#ifdef _OPENMP
        return threads > 0 ? threads : omp_get_max_threads();
#else
        return 1;
#endif
    }

    size_t id3(uint32_t x, uint32_t y, uint32_t z) const { return x + size_t(nx) * (y + size_t(ny) * z); }
    bool interior(uint32_t x, uint32_t y, uint32_t z) const {
        return x >= 1 && x + 2 <= nx && y >= 1 && y + 2 <= ny && z >= 1 && z + 2 <= nz;
    }
    T store(T v) const { return quantize<T>(p.pop_format, v); }

    // -- init (initialConditions.cu:33-140) --------------------------------
    void init() override {
        flush_denormals();
        for (auto &v : fld) std::fill(v.begin(), v.end(), T(0));
        std::fill(f.begin(), f.end(), T(0));
        std::fill(g.begin(), g.end(), T(0));
        if (p.water.enabled) {  // setWaterJet :33-56
            for (uint32_t z = 0; z < nz; ++z)
                for (uint32_t x = 0; x < nx; ++x) {
                    const T dx = T(x) - T(p.water.center_a), dz = T(z) - T(p.water.center_b);
                    if (dx * dx + dz * dz > T(p.water.r2)) continue;
                    fld[PHI][id3(x, 0, z)] = T(p.water.phi);
                    fld[UY][id3(x, 0, z)] = T(p.water.u);
                }
        }
        if (p.oil.enabled) {  // setOilJet :58-81
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t x = 0; x < nx; ++x) {
                    const T dx = T(x) - T(p.oil.center_a), dy = T(y) - T(p.oil.center_b);
                    if (dx * dx + dy * dy > T(p.oil.r2)) continue;
                    fld[PHI][id3(x, y, 0)] = T(p.oil.phi);
                    fld[UZ][id3(x, y, 0)] = T(p.oil.u);
                }
        }
        std::fill(fld[RHO].begin(), fld[RHO].end(), T(1));  // setInitialDensity :83-97
        // setDistros :99-140 (own cell, not pushed)
#pragma omp parallel for num_threads(nthreads())
        for (long long c = 0; c < (long long)N; ++c) {
            const T ux = fld[UX][c], uy = fld[UY][c], uz = fld[UZ][c];
            const T uu = T(1.5) * (ux * ux + uy * uy + uz * uz);
            for (int i = 0; i < q; ++i) {
                const T cu = L.as2 * (T(C27[i][0]) * ux + T(C27[i][1]) * uy + T(C27[i][2]) * uz);
                f[size_t(i) * N + c] = store(L.feq(i, fld[RHO][c], uu, cu));
            }
            for (int k = 0; k < 7; ++k) g[size_t(k) * N + c] = L.geq(k, fld[PHI][c], ux, uy, uz);
        }
    }

    // -- K1 phase::computePhase (phaseField.cuh:31-52) ---------------------
    void k_phase() override {
#pragma omp parallel for num_threads(nthreads())
        for (long long z = 1; z < (long long)nz - 1; ++z)
            for (uint32_t y = 1; y + 1 < ny; ++y)
                for (uint32_t x = 1; x + 1 < nx; ++x) {
                    const size_t c = id3(x, y, uint32_t(z));
                    T s = T(0);
                    for (int k = 0; k < 7; ++k) s += g[size_t(k) * N + c];
                    fld[PHI][c] = s;
                }
    }

    // Isotropic gradient with the hydrodynamic weights: 3 * sum_i w_i c_i v(x+c_i)
    // (phaseField.cuh:158-218; lbm.cu:157-217). `get` returns the stencil value.
    template <class Get>
    void grad(uint32_t x, uint32_t y, uint32_t z, Get get, T &gx, T &gy, T &gz) const {
        T sx = T(0), sy = T(0), sz = T(0);
        for (int i = 1; i < q; ++i) {
            const int cx = C27[i][0], cy = C27[i][1], cz = C27[i][2];
            const T v = L.w[i] * get(long(x) + cx, long(y) + cy, long(z) + cz);
            sx += T(cx) * v;
            sy += T(cy) * v;
            sz += T(cz) * v;
        }
        gx = L.as2 * sx;
        gy = L.as2 * sy;
        gz = L.as2 * sz;
    }

    // -- K2 phase::computeForces (phaseField.cuh:54-294) -------------------
    // phi-hat: phi at interior cells, 0 on the ghost layer and outside
    // (:107-143); normals and |grad| of phi-hat on every in-domain cell
    // (:147-227); curvature and F = -sigma*kappa*|grad|*n on interior cells,
    // F = 0 on the ghost layer (:231-293).
    void k_forces() override {
        const std::vector<T> &phi = fld[PHI];
        auto phat = [&](long x, long y, long z) -> T {
            if (x < 1 || y < 1 || z < 1 || x > long(nx) - 2 || y > long(ny) - 2 || z > long(nz) - 2) return T(0);
            return phi[id3(uint32_t(x), uint32_t(y), uint32_t(z))];
        };
#pragma omp parallel for num_threads(nthreads())
        for (long long z = 0; z < (long long)nz; ++z)
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t x = 0; x < nx; ++x) {
                    T gx, gy, gz;
                    grad(x, y, uint32_t(z), phat, gx, gy, gz);
                    const T ind = std::sqrt(gx * gx + gy * gy + gz * gz);
                    const T inv = T(1) / (ind + T(1e-9));
                    const size_t c = id3(x, y, uint32_t(z));
                    nnx[c] = gx * inv;
                    nny[c] = gy * inv;
                    nnz[c] = gz * inv;
                    nind[c] = ind;
                }
#pragma omp parallel for num_threads(nthreads())
        for (long long z = 0; z < (long long)nz; ++z)
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t x = 0; x < nx; ++x) {
                    const size_t c = id3(x, y, uint32_t(z));
                    if (!interior(x, y, uint32_t(z))) {
                        fld[FSX][c] = fld[FSY][c] = fld[FSZ][c] = T(0);
                        continue;
                    }
                    T s = T(0);
                    for (int i = 1; i < q; ++i) {
                        const size_t n = id3(x + C27[i][0], y + C27[i][1], uint32_t(z + C27[i][2]));
                        s += L.w[i] * (T(C27[i][0]) * nnx[n] + T(C27[i][1]) * nny[n] + T(C27[i][2]) * nnz[n]);
                    }
                    const T curvature = L.as2 * s;
                    const T stCurv = -sigma * curvature * nind[c];
                    fld[FSX][c] = stCurv * nnx[c];
                    fld[FSY][c] = stCurv * nny[c];
                    fld[FSZ][c] = stCurv * nnz[c];
                }
    }

    // -- K3 lbm::computeMoments (lbm.cu:33-128) ---------------------------
    void k_moments() override {
#pragma omp parallel for num_threads(nthreads())
        for (long long z = 1; z < (long long)nz - 1; ++z)
            for (uint32_t y = 1; y + 1 < ny; ++y)
                for (uint32_t x = 1; x + 1 < nx; ++x) {
                    const size_t c = id3(x, y, uint32_t(z));
                    T pop[27];
                    T rho = T(0);
                    for (int i = 0; i < q; ++i) {
                        pop[i] = f[size_t(i) * N + c];
                        rho += pop[i];
                    }
                    rho += T(1);
                    const T fx = fld[FSX][c], fy = fld[FSY][c], fz = fld[FSZ][c];
                    const T invRho = T(1) / rho;
                    T jx = T(0), jy = T(0), jz = T(0);
                    for (int i = 1; i < q; ++i) {
                        jx += T(C27[i][0]) * pop[i];
                        jy += T(C27[i][1]) * pop[i];
                        jz += T(C27[i][2]) * pop[i];
                    }
                    T ux = invRho * jx, uy = invRho * jy, uz = invRho * jz;
                    ux += fx * T(0.5) * invRho;
                    uy += fy * T(0.5) * invRho;
                    uz += fz * T(0.5) * invRho;
                    const T uu = T(1.5) * (ux * ux + uy * uy + uz * uz);
                    T pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
                    for (int i = 0; i < q; ++i) {
                        const T cx = T(C27[i][0]), cy = T(C27[i][1]), cz = T(C27[i][2]);
                        const T cu = L.as2 * (cx * ux + cy * uy + cz * uz);
                        const T fneq = pop[i] - L.feq(i, rho, uu, cu) + T(0.5) * L.force(i, cu, ux, uy, uz, fx, fy, fz);
                        pxx += fneq * cx * cx;
                        pyy += fneq * cy * cy;
                        pzz += fneq * cz * cz;
                        pxy += fneq * cx * cy;
                        pxz += fneq * cx * cz;
                        pyz += fneq * cy * cz;
                    }
                    const T trace = pxx + pyy + pzz;
                    pxx -= L.cs2 * trace;
                    pyy -= L.cs2 * trace;
                    pzz -= L.cs2 * trace;
                    fld[RHO][c] = rho;
                    fld[UX][c] = ux;
                    fld[UY][c] = uy;
                    fld[UZ][c] = uz;
                    fld[PXX][c] = pxx;
                    fld[PYY][c] = pyy;
                    fld[PZZ][c] = pzz;
                    fld[PXY][c] = pxy;
                    fld[PXZ][c] = pxz;
                    fld[PYZ][c] = pyz;
                }
    }

    // -- K4 lbm::streamCollide (lbm.cu:130-270) ----------------------------
    void k_collide() override {
        const std::vector<T> &phi = fld[PHI];
        auto real = [&](long x, long y, long z) -> T { return phi[id3(uint32_t(x), uint32_t(y), uint32_t(z))]; };
#pragma omp parallel for num_threads(nthreads())
        for (long long zz = 1; zz < (long long)nz - 1; ++zz)
            for (uint32_t y = 1; y + 1 < ny; ++y)
                for (uint32_t x = 1; x + 1 < nx; ++x) {
                    const uint32_t z = uint32_t(zz);
                    const size_t c = id3(x, y, z);
                    const T rho = fld[RHO][c], ux = fld[UX][c], uy = fld[UY][c], uz = fld[UZ][c];
                    const T pxx = fld[PXX][c], pyy = fld[PYY][c], pzz = fld[PZZ][c];
                    const T pxy = fld[PXY][c], pxz = fld[PXZ][c], pyz = fld[PYZ][c];
                    const T fx = fld[FSX][c], fy = fld[FSY][c], fz = fld[FSZ][c];
                    T gx, gy, gz;
                    grad(x, y, z, real, gx, gy, gz);
                    const T ind = std::sqrt(gx * gx + gy * gy + gz * gz);
                    const T invInd = T(1) / (ind + T(1e-9));
                    const T normx = gx * invInd, normy = gy * invInd, normz = gz * invInd;
                    const T ph = phi[c];
                    const T nu = std::fma(ph, visc_diff, visc_water);
                    T omco;
                    if (p.closure == SSMD_CLOSURE_NONE) {
                        const T omega = T(1) / (T(0.5) + L.as2 * nu);
                        omco = T(1) - omega;
                    } else {
                        omco = T(1) - T(1) / closure_tau(c, rho, pxx, pyy, pzz, pxy, pxz, pyz, ph,
                                                         T(0.5) + L.as2 * nu);
                    }
                    const T uu = T(1.5) * (ux * ux + uy * uy + uz * uz);
                    for (int i = 0; i < q; ++i) {
                        const T cx = T(C27[i][0]), cy = T(C27[i][1]), cz = T(C27[i][2]);
                        const T cu = L.as2 * (cx * ux + cy * uy + cz * uz);
                        const T feq = L.feq(i, rho, uu, cu);
                        const T force = L.force(i, cu, ux, uy, uz, fx, fy, fz);
                        const T fneq = L.fneq(i, pxx, pyy, pzz, pxy, pxz, pyz, ux, uy, uz);
                        const size_t dst = id3(x + C27[i][0], y + C27[i][1], z + C27[i][2]);
                        f[size_t(i) * N + dst] = store(feq + omco * fneq + T(0.5) * force);
                    }
                    T sharp = gamma * ph * (T(1) - ph);
                    // LEARNED flux closure (no reference counterpart): J = -D_t grad(phi)
                    // with D_t = max(flux_coeff tau_t / 3, -0.1) adds 4 J to the first
                    // moments that antidiff() carries as sharp * n
                    if (p.closure == SSMD_CLOSURE_LEARNED && p.flux_coeff != 0.0)
                        sharp -= T(4) * std::max(flux3 * fld[TAU_T][c], T(-0.1)) * ind;
                    for (int k = 0; k < 7; ++k) {
                        const T geq = L.geq(k, ph, ux, uy, uz);
                        const T hi = L.antidiff(k, sharp, normx, normy, normz);
                        const size_t dst = id3(x + C27[k][0], y + C27[k][1], z + C27[k][2]);
                        g[size_t(k) * N + dst] = geq + hi;
                    }
                }
    }

    // Subgrid closures (no reference counterpart; the GPU kernel's
    // closure_tau).  SMAGORINSKY: Hou et al. (1996), tau = (tau0 +
    // sqrt(tau0^2 + 18 Cs^2 sqrt(2 Pi:Pi) / rho)) / 2.  LEARNED: tau = tau0 +
    // tau_t; records the network inputs (strain rate S = -3 Pi / (2 rho tau)
    // times strain_scale, phi, |grad phi-hat| times grad_scale), unrounded.
    T closure_tau(size_t c, T rho, T pxx, T pyy, T pzz, T pxy, T pxz, T pyz, T ph, T tau0) {
        if (p.closure == SSMD_CLOSURE_SMAGORINSKY) return smag_tau<T>(tau0, pxx, pyy, pzz, pxy, pxz, pyz, rho, T(p.cs));
        const T tau = tau0 + fld[TAU_T][c];
        const T sc = T(-1.5) * T(p.strain_scale) / (rho * tau);
        const T v[8] = {pxx * sc, pyy * sc, pzz * sc, pxy * sc, pxz * sc, pyz * sc, ph, nind[c] * T(p.grad_scale)};
        for (int j = 0; j < 8; ++j) fld[FEAT0 + j][c] = v[j];
        return tau;
    }

    // -- boundaries (BoundaryConditions.cuh) -------------------------------
    bool in_dom(long x, long y, long z) const {
        return x >= 0 && y >= 0 && z >= 0 && x < long(nx) && y < long(ny) && z < long(nz);
    }

    // Inlet velocity with the seeded noise; `axis` is the face normal.
    void inlet_u(const ssmd_jet &j, int axis, uint64_t step, size_t cell, T &ux, T &uy, T &uz) const {
        const T u = T(j.u);
        T du[3] = {T(0), T(0), T(0)};
        if (p.perturb != 0.0) {
            const T amp = T(p.perturb) * u;
            for (int a = 0; a < 3; ++a) du[a] = amp * T(noise_uniform(p.seed, step, cell, uint64_t(a)));
        }
        ux = du[0];
        uy = du[1];
        uz = du[2];
        if (axis == 1) uy += u;
        else uz += u;
    }

    // applyWaterInflow (:36-99) on y = 0, and applyOilInflow (:101-168) on
    // z = 0.  The water inlet has no ghost-column check in the reference;
    // here an arrival outside the domain is skipped (the reference would
    // index out of bounds there, which its geometry never reaches).
    void inflow(const ssmd_jet &j, int axis, uint64_t step) {
        const uint32_t na = nx, nb = axis == 1 ? nz : ny;
        for (uint32_t b = 0; b < nb; ++b)
            for (uint32_t a = 0; a < na; ++a) {
                const T da = T(a) - T(j.center_a), db = T(b) - T(j.center_b);
                if (da * da + db * db > T(j.r2)) continue;
                const uint32_t x = a, y = axis == 1 ? 0 : b, z = axis == 1 ? b : 0;
                const size_t bnd = id3(x, y, z);
                T ux, uy, uz;
                inlet_u(j, axis, step, bnd, ux, uy, uz);
                const T rho = T(1), ph = T(j.phi);
                fld[RHO][bnd] = rho;
                fld[PHI][bnd] = ph;
                fld[UX][bnd] = ux;
                fld[UY][bnd] = uy;
                fld[UZ][bnd] = uz;
                const T uu = T(1.5) * (ux * ux + uy * uy + uz * uz);
                const T omco = axis == 1 ? omco_water : omco_oil;
                for (int i = 0; i < q; ++i) {
                    if (C27[i][axis] != 1) continue;
                    const long xx = long(x) + C27[i][0], yy = long(y) + C27[i][1], zz = long(z) + C27[i][2];
                    if (axis == 2) {  // :145-148
                        if (xx <= 0 || xx >= long(nx) - 1 || yy <= 0 || yy >= long(ny) - 1) continue;
                    } else if (!in_dom(xx, yy, zz)) {
                        continue;
                    }
                    const size_t fl = id3(uint32_t(xx), uint32_t(yy), uint32_t(zz));
                    const T cu = L.as2 * (T(C27[i][0]) * ux + T(C27[i][1]) * uy + T(C27[i][2]) * uz);
                    const T feq = L.feq(i, rho, uu, cu);
                    const T fneq = L.fneq(i, fld[PXX][fl], fld[PYY][fl], fld[PZZ][fl], fld[PXY][fl], fld[PXZ][fl],
                                          fld[PYZ][fl], fld[UX][fl], fld[UY][fl], fld[UZ][fl]);
                    f[size_t(i) * N + fl] = store(feq + omco * fneq);
                }
                const int k = axis == 1 ? 3 : 5;
                const T un = axis == 1 ? uy : uz;
                g[size_t(k) * N + id3(x + (axis == 0), y + (axis == 1), z + (axis == 2))] =
                    L.w7[k] * ph * (T(1) + T(4) * un);
            }
    }

    // applyOutflowY (:170-233) and applyOutflowZ (:235-293).
    void outflow(int axis) {
        const uint32_t nb = axis == 1 ? nz : ny;
        const uint32_t nn = axis == 1 ? ny : nz;
        for (uint32_t b = 0; b < nb; ++b)
            for (uint32_t a = 0; a < nx; ++a) {
                if (a == 0 || a == nx - 1 || b == 0 || b == nb - 1) continue;
                const uint32_t x = a, y = axis == 1 ? ny - 1 : b, z = axis == 1 ? b : nz - 1;
                const size_t bnd = id3(x, y, z);
                const size_t in = axis == 1 ? id3(x, ny - 2, z) : id3(x, y, nz - 2);
                for (int fid : {RHO, PHI, UX, UY, UZ}) fld[fid][bnd] = fld[fid][in];
                const T rho = fld[RHO][bnd], ph = fld[PHI][bnd];
                const T ux = fld[UX][bnd], uy = fld[UY][bnd], uz = fld[UZ][bnd];
                const T uu = T(1.5) * (ux * ux + uy * uy + uz * uz);
                for (int i = 0; i < q; ++i) {
                    if (C27[i][axis] != -1) continue;
                    const long xx = long(x) + C27[i][0];
                    const long yy = axis == 1 ? long(ny) - 2 : long(y) + C27[i][1];
                    const long zz = axis == 1 ? long(z) + C27[i][2] : long(nz) - 2;
                    if (axis == 1) {  // :210-213
                        if (xx <= 0 || xx >= long(nx) - 1 || zz <= 0 || zz >= long(nz) - 1) continue;
                    }
                    const size_t fl = id3(uint32_t(xx), uint32_t(yy), uint32_t(zz));
                    const T cu = L.as2 * (T(C27[i][0]) * ux + T(C27[i][1]) * uy + T(C27[i][2]) * uz);
                    const T feq = L.feq(i, rho, uu, cu);
                    const T fneq = L.fneq(i, fld[PXX][fl], fld[PYY][fl], fld[PZZ][fl], fld[PXY][fl], fld[PXZ][fl],
                                          fld[PYZ][fl], fld[UX][fl], fld[UY][fl], fld[UZ][fl]);
                    f[size_t(i) * N + fl] = store(feq + omco_ref * fneq);
                }
                const int k = axis == 1 ? 4 : 6;
                const T un = axis == 1 ? uy : uz;
                g[size_t(k) * N + in] = L.w7[k] * ph * (T(1) - T(4) * un);
                (void)nn;
            }
    }

    // periodicX (:295-349) and its mirrors.  mode REFERENCE exchanges the
    // arrival planes 1 and n-2; TRUE copies the ghost arrivals (pushed out
    // of the domain) back in with period n-2.  Ghost phi/u wrap is shared.
    void periodic(int axis, int mode) {
        const uint32_t n[3] = {nx, ny, nz};
        const int ax1 = axis == 0 ? 1 : 0, ax2 = axis == 2 ? 1 : 2;
        for (uint32_t b = 0; b < n[ax2]; ++b)
            for (uint32_t a = 0; a < n[ax1]; ++a) {
                auto at = [&](uint32_t s) {
                    uint32_t r[3];
                    r[axis] = s;
                    r[ax1] = a;
                    r[ax2] = b;
                    return id3(r[0], r[1], r[2]);
                };
                const size_t bL = at(1), bR = at(n[axis] - 2), gL = at(0), gR = at(n[axis] - 1);
                const size_t srcP = mode == SSMD_WRAP_TRUE ? gR : bR;  // feeds c_axis > 0 at plane 1
                const size_t srcM = mode == SSMD_WRAP_TRUE ? gL : bL;  // feeds c_axis < 0 at plane n-2
                for (int i = 0; i < q; ++i) {
                    if (C27[i][axis] > 0) f[size_t(i) * N + bL] = f[size_t(i) * N + srcP];
                    if (C27[i][axis] < 0) f[size_t(i) * N + bR] = f[size_t(i) * N + srcM];
                }
                for (int k = 1; k < 7; ++k) {
                    if (C27[k][axis] > 0) g[size_t(k) * N + bL] = g[size_t(k) * N + srcP];
                    if (C27[k][axis] < 0) g[size_t(k) * N + bR] = g[size_t(k) * N + srcM];
                }
                for (int fid : {PHI, UX, UY, UZ}) {
                    fld[fid][gL] = fld[fid][bR];
                    fld[fid][gR] = fld[fid][bL];
                }
            }
    }

    void k_boundaries(uint64_t step) override {
        if (p.water.enabled) inflow(p.water, 1, step);
        if (p.oil.enabled) inflow(p.oil, 2, step);
        if (p.outflow_y) outflow(1);
        if (p.outflow_z) outflow(2);
        if (p.wrap_x) periodic(0, p.wrap_x);
        if (p.wrap_y) periodic(1, p.wrap_y);
        if (p.wrap_z) periodic(2, p.wrap_z);
    }

    // -- accessors ----------------------------------------------------------
    double *field(int) override { return nullptr; }
    int get_field(int id, double *out) override {
        if (id < 0 || id >= NFIELDS) return SSMD_EINVAL;
        for (size_t c = 0; c < N; ++c) out[c] = double(fld[id][c]);
        return 0;
    }
    int set_field(int id, const double *in) override {
        if (id < 0 || id >= NFIELDS) return SSMD_EINVAL;
        for (size_t c = 0; c < N; ++c) fld[id][c] = T(in[c]);
        return 0;
    }
    int get_pops(int which, double *out) override {
        const std::vector<T> &v = which == 0 ? f : g;
        for (size_t c = 0; c < v.size(); ++c) out[c] = double(v[c]);
        return 0;
    }
    int set_pops(int which, const double *in) override {
        std::vector<T> &v = which == 0 ? f : g;
        for (size_t c = 0; c < v.size(); ++c) v[c] = which == 0 ? store(T(in[c])) : T(in[c]);
        return 0;
    }
    void sums(double *out) override {
        double s[4] = {0, 0, 0, 0};
        for (uint32_t z = 1; z + 1 < nz; ++z)
            for (uint32_t y = 1; y + 1 < ny; ++y)
                for (uint32_t x = 1; x + 1 < nx; ++x) {
                    const size_t c = id3(x, y, z);
                    const double r = double(fld[RHO][c]);
                    s[0] += double(fld[PHI][c]);
                    s[1] += r * double(fld[UX][c]);
                    s[2] += r * double(fld[UY][c]);
                    s[3] += r * double(fld[UZ][c]);
                }
        for (int k = 0; k < 4; ++k) out[k] = s[k];
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// C API used by tests/ (ctypes) and by bench.py's cpu_baseline leg.
extern "C" {

struct oracle_handle {
    std::unique_ptr<Base> impl;
};

size_t oracle_params_size(void) { return sizeof(ssmd_params); }

oracle_handle *oracle_create(const ssmd_params *p, int fp64) {
    if (!p || (p->q != 19 && p->q != 27) || p->nx < 3 || p->ny < 3 || p->nz < 3) return nullptr;
    auto *h = new oracle_handle;
    if (fp64)
        h->impl.reset(new Oracle<double>(*p));
    else
        h->impl.reset(new Oracle<float>(*p));
    return h;
}
void oracle_destroy(oracle_handle *h) { delete h; }
void oracle_set_threads(oracle_handle *h, int n) { h->impl->set_threads(n); }
void oracle_init(oracle_handle *h) { h->impl->init(); }
void oracle_step(oracle_handle *h, uint64_t step) { h->impl->step(step); }
void oracle_run(oracle_handle *h, uint64_t step0, uint64_t nsteps) {
    for (uint64_t s = 0; s < nsteps; ++s) h->impl->step(step0 + s);
}
void oracle_k_phase(oracle_handle *h) { h->impl->k_phase(); }
void oracle_k_forces(oracle_handle *h) { h->impl->k_forces(); }
void oracle_k_moments(oracle_handle *h) { h->impl->k_moments(); }
void oracle_k_collide(oracle_handle *h) { h->impl->k_collide(); }
void oracle_k_boundaries(oracle_handle *h, uint64_t step) { h->impl->k_boundaries(step); }
int oracle_get_field(oracle_handle *h, int id, double *out) { return h->impl->get_field(id, out); }
int oracle_set_field(oracle_handle *h, int id, const double *in) { return h->impl->set_field(id, in); }
int oracle_get_pops(oracle_handle *h, int which, double *out) { return h->impl->get_pops(which, out); }
int oracle_set_pops(oracle_handle *h, int which, const double *in) { return h->impl->set_pops(which, in); }
void oracle_sums(oracle_handle *h, double *out4) { h->impl->sums(out4); }

// Pointwise lattice functions in FP64, for the invariant tests.
double oracle_w(int q, int i) { return q == 7 ? w_phase(i) : w_hydro(q, i); }
int oracle_c(int i, int axis) { return C27[i][axis]; }
double oracle_feq(int q, int i, double rho, double ux, double uy, double uz) {
    Lat<double> L(q);
    const double uu = 1.5 * (ux * ux + uy * uy + uz * uz);
    const double cu = 3.0 * (C27[i][0] * ux + C27[i][1] * uy + C27[i][2] * uz);
    return L.feq(i, rho, uu, cu);
}
double oracle_fneq(int q, int i, const double *pi6, double ux, double uy, double uz) {
    Lat<double> L(q);
    return L.fneq(i, pi6[0], pi6[1], pi6[2], pi6[3], pi6[4], pi6[5], ux, uy, uz);
}
double oracle_force(int q, int i, double ux, double uy, double uz, double fx, double fy, double fz) {
    Lat<double> L(q);
    const double cu = 3.0 * (C27[i][0] * ux + C27[i][1] * uy + C27[i][2] * uz);
    return L.force(i, cu, ux, uy, uz, fx, fy, fz);
}
double oracle_geq(int k, double phi, double ux, double uy, double uz) {
    Lat<double> L(19);
    return L.geq(k, phi, ux, uy, uz);
}
double oracle_antidiff(int k, double sharp, double nx, double ny, double nz) {
    Lat<double> L(19);
    return L.antidiff(k, sharp, nx, ny, nz);
}
double oracle_smag_tau(double tau0, const double *pi6, double rho, double cs) {
    return smag_tau<double>(tau0, pi6[0], pi6[1], pi6[2], pi6[3], pi6[4], pi6[5], rho, cs);
}
uint16_t oracle_f32_to_f16(float v) { return f32_to_f16(v); }
float oracle_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
uint16_t oracle_f32_to_bf16(float v) { return f32_to_bf16(v); }
float oracle_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }
uint64_t oracle_splitmix64(uint64_t z) { return splitmix64(z); }
float oracle_noise(uint64_t seed, uint64_t step, uint64_t cell, uint64_t comp) {
    return noise_uniform(seed, step, cell, comp);
}

}  // extern "C"
