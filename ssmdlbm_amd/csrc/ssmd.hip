// ssmd.hip -- C ABI of the gfx950 hot path (declared in include/ssmd.h).
//
// Host responsibilities mirrored from the reference driver:
//   FieldAllocate (fieldAllocate/FieldAllocate.cuh:42-162)  -> ssmd_alloc/free
//   init kernels (main.cu:109-112)                           -> ssmd_init
//   graph + boundary launches per step (main.cu:159-172)     -> ssmd_step
//   graph::captureGraph (cuda/CUDAGraph.cuh:34-57)           -> ssmd_graph_*
// Errors are returned, never aborted on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/ssmd.h"
#define SSMD_STEP_KERNELS 1  // the non-template kernels live in this translation unit
#include "launch.hpp"

using namespace ssmd;

namespace {

int herr(hipError_t e) { return e == hipSuccess ? 0 : int(e); }

#define SSMD_TRY(...)                          \
    do {                                       \
        const int _rc = (__VA_ARGS__);         \
        if (_rc != 0) return _rc;              \
    } while (0)
#define SSMD_HIP(...) SSMD_TRY(herr(__VA_ARGS__))

size_t pop_size(int fmt) { return fmt == SSMD_POP_F32 ? 4 : 2; }

KParams make_kparams(const ssmd_params &p) {
    KParams k{};
    k.nx = int(p.nx);
    k.ny = int(p.ny);
    k.nz = int(p.nz);
    k.sxy = p.nx * p.ny;
    k.N = p.nx * p.ny * p.nz;
    // relaxation:: (globalFunctions.cuh:232-283): formed in double, cast.
    k.visc_water = float(p.visc_water);
    k.visc_diff = float(p.visc_oil) - float(p.visc_water);
    auto omega = [](double nu) { return 1.0 / (0.5 + 3.0 * nu); };
    k.omco_water = 1.0f - float(omega(p.visc_water));
    k.omco_oil = 1.0f - float(omega(p.visc_oil));
    k.omco_ref = k.omco_water;
    k.sigma = float(p.sigma);
    k.gamma = float(p.gamma);
    k.perturb = float(p.perturb);
    k.seed = p.seed;
    auto jet = [](const ssmd_jet &j) {
        JetK r{};
        r.enabled = j.enabled;
        r.ca = float(j.center_a);
        r.cb = float(j.center_b);
        r.r2 = float(j.r2);
        r.u = float(j.u);
        r.phi = float(j.phi);
        return r;
    };
    k.water = jet(p.water);
    k.oil = jet(p.oil);
    k.wrap_x = p.wrap_x;
    k.wrap_y = p.wrap_y;
    k.wrap_z = p.wrap_z;
    k.bc_planes = (p.water.enabled ? 1 : 0) | (p.outflow_y ? 2 : 0) | (p.oil.enabled ? 4 : 0) | (p.outflow_z ? 8 : 0);
    // g in moment form is 16 B per cell; cached stores measured +1.8% on an
    // 8.4 M-cell box (134 MB of g) and -1% on a 67 M-cell one (1.07 GB)
    k.g_cached = uint64_t(p.nx) * p.ny * p.nz * 16u <= (uint64_t(160) << 20) ? 1 : 0;
    k.nt_pull = k.g_cached;
    k.closure = p.closure;
    k.smag = float(18.0 * p.cs * p.cs);
    k.strain_scale = float(p.strain_scale);
    k.grad_scale = float(p.grad_scale);
    k.flux3 = float(p.flux_coeff / 3.0);
    return k;
}

Macro make_macro(const ssmd_fields &d) {
    Macro m{};
    m.rho = d.rho;
    m.ux = d.ux;
    m.uy = d.uy;
    m.uz = d.uz;
    m.pxx = d.pxx;
    m.pyy = d.pyy;
    m.pzz = d.pzz;
    m.pxy = d.pxy;
    m.pxz = d.pxz;
    m.pyz = d.pyz;
    m.fsx = d.fsx;
    m.fsy = d.fsy;
    m.fsz = d.fsz;
    m.tau_t = d.tau_t;
    m.feat = d.feat;
    return m;
}

bool fields_ready(const ssmd_params &p, const ssmd_fields *d) {
    if (!(d && d->rho && d->ux && d->uy && d->uz && d->pxx && d->pyy && d->pzz && d->pxy && d->pxz && d->pyz &&
          d->phi && d->fsx && d->fsy && d->fsz && d->f && d->g && d->f_next && d->g_next))
        return false;
    if (p.closure == SSMD_CLOSURE_LEARNED)
        return d->tau_t && d->feat && d->mlp && (reinterpret_cast<uintptr_t>(d->feat) & 15) == 0;
    return true;
}

int tile_of(const ssmd_params &p) { return (p.tile >= 0 && p.tile < kNumTiles) ? p.tile : 0; }

// Calls fn.template operator()<Q, P>() for the lattice and storage format.
template <class Fn>
int dispatch(const ssmd_params &p, Fn &&fn) {
    switch (p.q * 10 + p.pop_format) {
        case 190: return fn.template operator()<19, PopF32>();
        case 191: return fn.template operator()<19, PopF16>();
        case 192: return fn.template operator()<19, PopBF16>();
        case 270: return fn.template operator()<27, PopF32>();
        case 271: return fn.template operator()<27, PopF16>();
        case 272: return fn.template operator()<27, PopBF16>();
        default: return SSMD_EUNSUPPORTED;
    }
}

// The face operations of a parameter block, in the order a step runs them.
int face_kinds(const ssmd_params &p, int *kinds) {
    int n = 0;
    if (p.water.enabled) kinds[n++] = kInflowY;
    if (p.oil.enabled) kinds[n++] = kInflowZ;
    if (p.outflow_y) kinds[n++] = kOutflowY;
    if (p.outflow_z) kinds[n++] = kOutflowZ;
    // the two wrap modes differ only in the population copies, which the pull
    // addresses carry; the ghost phi/u wrap of x then y is one pass
    if (p.wrap_x && p.wrap_y) {
        kinds[n++] = kPeriodicXY;
    } else {
        if (p.wrap_x) kinds[n++] = p.wrap_x == SSMD_WRAP_TRUE ? kPeriodicXTrue : kPeriodicXRef;
        if (p.wrap_y) kinds[n++] = p.wrap_y == SSMD_WRAP_TRUE ? kPeriodicYTrue : kPeriodicYRef;
    }
    if (p.wrap_z) kinds[n++] = p.wrap_z == SSMD_WRAP_TRUE ? kPeriodicZTrue : kPeriodicZRef;
    return n;
}

bool is_inflow(int kind) { return kind == kInflowY || kind == kInflowZ; }

// Runtime face kind -> template instantiation.
template <class Fn>
void face_dispatch(int kind, Fn &&fn) {
    switch (kind) {
        case kInflowY: fn.template operator()<kInflowY>(); break;
        case kInflowZ: fn.template operator()<kInflowZ>(); break;
        case kOutflowY: fn.template operator()<kOutflowY>(); break;
        case kOutflowZ: fn.template operator()<kOutflowZ>(); break;
        case kPeriodicXRef: fn.template operator()<kPeriodicXRef>(); break;
        case kPeriodicXTrue: fn.template operator()<kPeriodicXTrue>(); break;
        case kPeriodicYRef: fn.template operator()<kPeriodicYRef>(); break;
        case kPeriodicYTrue: fn.template operator()<kPeriodicYTrue>(); break;
        case kPeriodicZRef: fn.template operator()<kPeriodicZRef>(); break;
        case kPeriodicZTrue: fn.template operator()<kPeriodicZTrue>(); break;
        case kPeriodicXY: fn.template operator()<kPeriodicXY>(); break;
        default: break;
    }
}

int face_axis_host(int kind) { return face_axis(kind); }

// `counter` (graph steps): the device step index, which the phase kernel
// advances at the start of each step (see launch_boundaries).
int launch_phase(const ssmd_params &p, const ssmd_fields &d, hipStream_t s, uint64_t *counter = nullptr) {
    const KParams k = make_kparams(p);
    if (p.nx % 4 == 0) {
        // one block per 256 quads of interior rows and phase_quad_lz planes
        const unsigned quads = (p.nx / 4) * (p.ny - 2);
        const int lz = phase_quad_lz(uint64_t(p.nx) * p.ny * p.nz);
        const dim3 grid(((quads + 255) / 256) * unsigned((p.nz - 2 + lz - 1) / lz));
        if (lz == 16)
            hipLaunchKernelGGL(k_phase_quad<16>, grid, dim3(256), 0, s, k, d.g, d.phi, counter);
        else
            hipLaunchKernelGGL(k_phase_quad<4>, grid, dim3(256), 0, s, k, d.g, d.phi, counter);
    } else {
        const unsigned chunks = unsigned((p.nz - 2 + kPhaseLZ - 1) / kPhaseLZ);
        // one block per 64 x 4 tile of interior rows and kPhaseLZ planes (the kernel derives the tile)
        const unsigned tiles = (unsigned(p.nx + 63) / 64) * (unsigned(p.ny - 2 + 3) / 4) * chunks;
        hipLaunchKernelGGL(k_phase, dim3(tiles), dim3(64, 4), 0, s, k, d.g, d.phi, counter);
    }
    return herr(hipGetLastError());
}

LatticeArgs lattice_args(const ssmd_params &p, const ssmd_fields &d, int lz) {
    LatticeArgs a{};
    a.k = make_kparams(p);
    a.fin = d.f;
    a.fout = d.f_next;
    a.gout = d.g_next;
    a.phi = d.phi;
    a.mac = make_macro(d);
    a.lz = lz;
    a.xcd_order = 0;
    return a;
}

int launch_lattice_sel(const ssmd_params &p, const ssmd_fields &d, uint32_t flags, int tile, int lz, hipStream_t s) {
    const bool macros = (flags & SSMD_STEP_MACROS) != 0;
    // the macro-writing kernel is built for tile 0 only; per-cell results do not depend on the tile
    const int t = macros ? 0 : tile;
    LatticeArgs a = lattice_args(p, d, lz > 0 ? lz : auto_lz(int(p.nx), int(p.ny), int(p.nz), t));
    a.xcd_order = kTiles[t].xcd_order;
    return dispatch(p, [&]<int Q, class P>() { return herr(launch_lattice<Q, P>(t, macros, a, s)); });
}

int launch_lattice(const ssmd_params &p, const ssmd_fields &d, uint32_t flags, hipStream_t s) {
    return launch_lattice_sel(p, d, flags, tile_of(p), p.lz, s);
}

// The face kernels of a step, in order.  `counter` null: the inlet noise
// takes `step`.  Otherwise the inlets read the device step index there, and
//   graph = false: the last face kernel that does not read it (or k_bump)
//                  advances it, so a caller may run step after step;
//   graph = true:  the step's phase kernel has advanced it already (graphs
//                  launch the phase kernel with the counter), and on steps that
//                  write no ghost u the faces among {z inlet, z outflow, x-y
//                  ring} run as one launch (k_faces_z), the same results as
//                  the kernels in order.
int launch_boundaries(const ssmd_params &p, const ssmd_fields &d, uint64_t *counter, uint64_t step, uint32_t sflags,
                      hipStream_t s, bool graph = false) {
    const KParams k = make_kparams(p);
    const Macro m = make_macro(d);
    int kinds[8];
    const int nk = face_kinds(p, kinds);
    const bool macros = (sflags & SSMD_STEP_MACROS) != 0;
    if (graph && !macros && nk >= 2) {
        bool in = false, out = false, ring = false, other = false;
        for (int i = 0; i < nk; ++i) {
            in = in || kinds[i] == kInflowZ;
            out = out || kinds[i] == kOutflowZ;
            ring = ring || kinds[i] == kPeriodicXY;
            other = other || (kinds[i] != kInflowZ && kinds[i] != kOutflowZ && kinds[i] != kPeriodicXY);
        }
        if (!other) {
            return dispatch(p, [&]<int Q, class P>() {
                using PT = typename P::T;
                PT *fo = static_cast<PT *>(d.f_next);
                const uint32_t zface = uint32_t((k.nx + 63) / 64) * uint32_t((k.ny + 3) / 4);
                const uint32_t nin = in ? zface * uint32_t(crossing_count<Q>()) : 0u;
                const uint32_t nout = out ? zface * uint32_t(crossing_count<Q>()) : 0u;
                const uint32_t nring = ring ? uint32_t((k.ny + k.nx - 2 + 63) / 64) * uint32_t((k.nz + 3) / 4) : 0u;
                const dim3 grid(nin + nout + nring), block(64, 4);
                if (in && out && ring)
                    hipLaunchKernelGGL((k_faces_z<Q, P, true, true, true>), grid, block, 0, s, k, fo, d.g_next, d.phi,
                                       m, counter, nin, nout);
                else if (in && out)
                    hipLaunchKernelGGL((k_faces_z<Q, P, true, true, false>), grid, block, 0, s, k, fo, d.g_next,
                                       d.phi, m, counter, nin, nout);
                else if (in && ring)
                    hipLaunchKernelGGL((k_faces_z<Q, P, true, false, true>), grid, block, 0, s, k, fo, d.g_next,
                                       d.phi, m, counter, nin, nout);
                else
                    hipLaunchKernelGGL((k_faces_z<Q, P, false, true, true>), grid, block, 0, s, k, fo, d.g_next,
                                       d.phi, m, counter, nin, nout);
                return herr(hipGetLastError());
            });
        }
    }
    int last = -1;  // last kind that may advance the step counter
    if (!graph)
        for (int i = 0; i < nk; ++i)
            if (!is_inflow(kinds[i])) last = i;
    return dispatch(p, [&]<int Q, class P>() {
        using PT = typename P::T;
        PT *fo = static_cast<PT *>(d.f_next);
        for (int i = 0; i < nk; ++i) {
            int na, nb;
            face_extent(face_axis_host(kinds[i]), k.nx, k.ny, k.nz, na, nb);
            const dim3 grid((na + 63) / 64, (nb + 3) / 4, face_layers<Q>(kinds[i]));
            face_dispatch(kinds[i], [&]<int KIND>() {
                const int fl = int(i == last) | (macros ? 2 : 0);
                hipLaunchKernelGGL((k_face<Q, P, KIND>), grid, dim3(64, 4), 0, s, k, fo, d.g_next, d.phi, m, counter,
                                   step, fl);
            });
        }
        if (!graph && counter && last < 0 && nk > 0) hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, s, counter);
        return herr(hipGetLastError());
    });
}

// One step on the buffers as they are in *d (no swap).
// `graph`: a captured step, whose phase kernel advances the device step
// counter (launch_boundaries).
int step_dispatch(const ssmd_params &p, const ssmd_fields &d, uint64_t *counter, uint64_t step, uint32_t flags,
                  hipStream_t s, bool graph = false) {
    SSMD_TRY(launch_phase(p, d, s, graph ? counter : nullptr));
    SSMD_TRY(launch_lattice(p, d, flags, s));
    SSMD_TRY(launch_boundaries(p, d, counter, step, flags, s, graph));
    if (p.closure == SSMD_CLOSURE_LEARNED)
        return launch_closure_update(d.mlp, d.feat, d.tau_t, p.nx * p.ny * p.nz, float(p.tau_scale),
                                     float(p.tau_floor), s);
    return SSMD_OK;
}

void swap_buffers(ssmd_fields *d) {
    void *f = d->f;
    d->f = d->f_next;
    d->f_next = f;
    float *g = d->g;
    d->g = d->g_next;
    d->g_next = g;
}

}  // namespace

// Steps per block graph.  Consecutive graph launches on a stream are ~8 us
// apart on MI355X (2% of a 128x128x512 D3Q27 step) while the kernels inside
// one graph follow each other with no gap, so runs replay blocks of
// kGraphBlock steps and single steps only for the remainder.  Even, so a
// block ends on the buffer parity it started from.
constexpr int kGraphBlock = 16;

struct ssmd_graph {
    ssmd_params p;
    void *fa, *fb;  // buffer identities: parity 0 reads fa
    float *ga, *gb;
    hipGraphExec_t exec[2];   // one step from parity 0 / 1
    hipGraph_t graph[2];
    hipGraphExec_t block_exec[2];  // kGraphBlock steps from parity 0 / 1
    hipGraph_t block_graph[2];
    uint64_t *counter;
    // LEARNED closure: the network pass forks onto a second stream during
    // capture (default priority), joined before the next lattice launch
    hipStream_t side = nullptr;
    hipEvent_t lattice_done = nullptr, network_done = nullptr;
};

extern "C" {

int ssmd_abi_version(void) { return SSMD_ABI_VERSION; }

const char *ssmd_error_string(int code) {
    switch (code) {
        case SSMD_OK: return "success";
        case SSMD_EINVAL: return "invalid argument";
        case SSMD_EUNSUPPORTED: return "unsupported parameter combination";
        case SSMD_ENOTREADY: return "fields or graph not ready";
        default: return code > 0 ? hipGetErrorString(hipError_t(code)) : "unknown error";
    }
}

int ssmd_params_reference_default(ssmd_params *p, int32_t q) {
    if (!p || (q != 19 && q != 27)) return SSMD_EINVAL;
    std::memset(p, 0, sizeof(*p));
    // constants.cuh:70-101
    const uint32_t res = 128;
    p->nx = res;
    p->ny = res * 2;
    p->nz = res * 2;
    p->q = q;
    p->pop_format = SSMD_POP_F32;
    const float uw = 0.05f, uo = 0.05f;  // physics::u_water / u_oil are scalar_t constants
    const double u_water = double(uw), u_oil = double(uo);
    const int diam = 13, radius = diam / 2;
    p->visc_water = u_water * diam / 1400.0;
    p->visc_oil = u_oil * diam / 450.0;
    p->sigma = double((uo * uo * float(diam)) / 500.0f);
    // gamma = kappa / cs2, kappa = 4 * cs2_7 * (tau_g - 1/2) / width, tau_g = 1, width = 4
    const float diff_int = 0.25f * (1.0f - 0.5f);
    const float kappa = 4.0f * diff_int / 4.0f;
    p->gamma = double(kappa / 0.25f);
    // geometry:: (globalFunctions.cuh:170-206)
    const double cx = double(float(res - 1) * 0.5f);
    const double cy = double(float(p->ny - 1) * 0.5f);
    const double cz = double(float(p->nz - 1) * 0.5f);
    p->water.enabled = 1;
    p->water.center_a = cx;
    p->water.center_b = double(0.7f * float(cz));
    p->water.r2 = double(radius * radius);
    p->water.diameter = diam;
    p->water.u = u_water;
    p->water.phi = 0.0;
    p->oil.enabled = 1;
    p->oil.center_a = cx;
    p->oil.center_b = double(0.5f * float(cy));
    p->oil.r2 = double(radius * radius);
    p->oil.diameter = diam;
    p->oil.u = u_oil;
    p->oil.phi = 1.0;
    p->outflow_y = 1;
    p->outflow_z = 1;
    p->wrap_x = SSMD_WRAP_REFERENCE;
    p->wrap_y = SSMD_WRAP_NONE;
    p->wrap_z = SSMD_WRAP_NONE;
    return SSMD_OK;
}

int ssmd_validate(const ssmd_params *p) {
    if (!p) return SSMD_EINVAL;
    if (p->q != 19 && p->q != 27) return SSMD_EINVAL;
    if (p->pop_format < 0 || p->pop_format > 2) return SSMD_EINVAL;
    if (p->nx < 5 || p->ny < 5 || p->nz < 5) return SSMD_EINVAL;
    // The kernels address a population or phase plane with 32-bit byte
    // offsets (4-byte elements at most), one plane of slack for the pulls.
    const uint64_t n = uint64_t(p->nx) * p->ny * p->nz;
    if (4 * (n + uint64_t(p->nx) * p->ny) >= (uint64_t(1) << 32)) return SSMD_EUNSUPPORTED;
    // launch grids put rows and planes on grid y / z (at most 65535 blocks)
    if (p->nx > 65536 || p->ny > 65536 || p->nz > 65536) return SSMD_EUNSUPPORTED;
    for (int w : {p->wrap_x, p->wrap_y, p->wrap_z})
        if (w < SSMD_WRAP_NONE || w > SSMD_WRAP_TRUE) return SSMD_EINVAL;
    if (p->wrap_y && (p->water.enabled || p->outflow_y)) return SSMD_EINVAL;
    if (p->wrap_z && (p->oil.enabled || p->outflow_z)) return SSMD_EINVAL;
    if (!(p->visc_water > 0.0) || !(p->visc_oil > 0.0)) return SSMD_EINVAL;
    if (p->tile < 0 || p->tile >= kNumTiles || p->lz < 0) return SSMD_EINVAL;
    // Periodic wraps are folded into the pull addresses from the first step
    // on, so the t=0 state must already be periodic along a wrapped axis:
    // inlet discs (the only non-uniform part of it) keep clear of those faces.
    auto disc_clear = [](const ssmd_jet &j, double lo_hi_n, bool axis_a) {
        if (!j.enabled) return true;
        const double r = std::sqrt(j.r2) + 1.0;
        const double c = axis_a ? j.center_a : j.center_b;
        return c - r > 2.0 && c + r < lo_hi_n - 3.0;
    };
    if (p->wrap_x && !(disc_clear(p->water, p->nx, true) && disc_clear(p->oil, p->nx, true))) return SSMD_EUNSUPPORTED;
    if (p->wrap_y && !disc_clear(p->oil, p->ny, false)) return SSMD_EUNSUPPORTED;
    if (p->wrap_z && !disc_clear(p->water, p->nz, false)) return SSMD_EUNSUPPORTED;
    if (p->closure < SSMD_CLOSURE_NONE || p->closure > SSMD_CLOSURE_LEARNED) return SSMD_EINVAL;
    if (p->closure == SSMD_CLOSURE_SMAGORINSKY && !(p->cs >= 0.0)) return SSMD_EINVAL;
    if (p->closure == SSMD_CLOSURE_LEARNED &&
        !(p->tau_scale >= 0.0 && std::isfinite(p->strain_scale) && std::isfinite(p->grad_scale)))
        return SSMD_EINVAL;
    // tau = 1/2 + 3 nu(phi) + tau_t must stay above 1/2 with margin
    if (!(p->tau_floor <= 0.0 && p->tau_floor >= -0.9 * 3.0 * std::min(p->visc_water, p->visc_oil)))
        return SSMD_EINVAL;
    if (!std::isfinite(p->flux_coeff) || (p->flux_coeff != 0.0 && p->closure != SSMD_CLOSURE_LEARNED))
        return SSMD_EINVAL;
    return SSMD_OK;
}

size_t ssmd_bytes(const ssmd_params *p, int32_t which) {
    if (ssmd_validate(p) != 0) return 0;
    const size_t n = size_t(p->nx) * p->ny * p->nz;
    if (which == 0) return n * 4;
    if (which == 1) return n * size_t(p->q) * pop_size(p->pop_format);
    if (which == 2) return n * 4 * 4;
    return 0;
}

int ssmd_alloc(const ssmd_params *p, ssmd_fields *d) {
    SSMD_TRY(ssmd_validate(p));
    if (!d) return SSMD_EINVAL;
    std::memset(d, 0, sizeof(*d));
    const size_t sb = ssmd_bytes(p, 0), fb = ssmd_bytes(p, 1), gb = ssmd_bytes(p, 2);
    float **scal[] = {&d->rho, &d->ux, &d->uy, &d->uz, &d->pxx, &d->pyy, &d->pzz,
                      &d->pxy, &d->pxz, &d->pyz, &d->phi, &d->fsx, &d->fsy, &d->fsz};
    d->owned = 1;
    auto fail = [&](hipError_t e) {
        ssmd_free(d);
        return int(e);
    };
    for (float **q : scal) {
        hipError_t e = hipMalloc(reinterpret_cast<void **>(q), sb);
        if (e != hipSuccess) return fail(e);
        e = hipMemset(*q, 0, sb);
        if (e != hipSuccess) return fail(e);
    }
    void **pops[] = {&d->f, &d->f_next};
    for (void **q : pops) {
        hipError_t e = hipMalloc(q, fb);
        if (e != hipSuccess) return fail(e);
        e = hipMemset(*q, 0, fb);
        if (e != hipSuccess) return fail(e);
    }
    {
        // one g buffer: the phase update reads g before the lattice kernel of
        // the same step writes it (g_next), so the pair can share it
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&d->g), gb);
        if (e != hipSuccess) return fail(e);
        d->g_next = d->g;
        e = hipMemset(d->g, 0, gb);
        if (e != hipSuccess) return fail(e);
    }
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&d->step_counter), sizeof(uint64_t));
    if (e != hipSuccess) return fail(e);
    e = hipMemset(d->step_counter, 0, sizeof(uint64_t));
    if (e != hipSuccess) return fail(e);
    e = hipMalloc(reinterpret_cast<void **>(&d->reduce), SSMD_REDUCE_DOUBLES * sizeof(double));
    if (e != hipSuccess) return fail(e);
    if (p->closure == SSMD_CLOSURE_LEARNED) {
        // zero weights: the network outputs 0 (tau_t = 0) until ssmd_mlp_pack
        const size_t n = size_t(p->nx) * p->ny * p->nz;
        void **bufs[] = {reinterpret_cast<void **>(&d->tau_t), reinterpret_cast<void **>(&d->feat), &d->mlp};
        const size_t bytes[] = {n * sizeof(float), n * SSMD_MLP_INPUTS * sizeof(uint16_t), ssmd_mlp_packed_bytes()};
        for (int i = 0; i < 3; ++i) {
            e = hipMalloc(bufs[i], bytes[i]);
            if (e != hipSuccess) return fail(e);
            e = hipMemset(*bufs[i], 0, bytes[i]);
            if (e != hipSuccess) return fail(e);
        }
    }
    return SSMD_OK;
}

int ssmd_free(ssmd_fields *d) {
    if (!d) return SSMD_EINVAL;
    if (!d->owned) return SSMD_OK;
    void *ptrs[] = {d->rho, d->ux,  d->uy,  d->uz,  d->pxx, d->pyy, d->pzz,    d->pxy,         d->pxz,
                    d->pyz, d->phi, d->fsx, d->fsy, d->fsz, d->f,   d->f_next, d->g,           d->g_next,
                    d->step_counter, d->reduce, d->tau_t, d->feat, d->mlp};
    int rc = 0;
    const size_t np = sizeof(ptrs) / sizeof(ptrs[0]);
    for (size_t i = 0; i < np; ++i) {
        bool seen = false;  // g_next may be g itself: each buffer freed once
        for (size_t j = 0; j < i; ++j) seen = seen || ptrs[j] == ptrs[i];
        if (ptrs[i] && !seen) {
            const int r = herr(hipFree(ptrs[i]));
            if (r && !rc) rc = r;
        }
    }
    std::memset(d, 0, sizeof(*d));
    return rc;
}

int ssmd_init(const ssmd_params *p, ssmd_fields *d, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    const KParams k = make_kparams(*p);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid((k.N + 255) / 256), block(256);
    const Macro m = make_macro(*d);
    const int key = p->q * 10 + p->pop_format;
    switch (key) {
#define SSMD_INIT(KEY, Q, PT)                                                                                      \
    case KEY:                                                                                                      \
        hipLaunchKernelGGL((k_init<Q, PT>), grid, block, 0, s, k, static_cast<PT::T *>(d->f),                     \
                           static_cast<PT::T *>(d->f_next), d->g, d->g_next, d->phi, m);                           \
        break;
        SSMD_INIT(190, 19, PopF32)
        SSMD_INIT(191, 19, PopF16)
        SSMD_INIT(192, 19, PopBF16)
        SSMD_INIT(270, 27, PopF32)
        SSMD_INIT(271, 27, PopF16)
        SSMD_INIT(272, 27, PopBF16)
#undef SSMD_INIT
        default: return SSMD_EUNSUPPORTED;
    }
    return herr(hipGetLastError());
}

int ssmd_step(const ssmd_params *p, ssmd_fields *d, uint64_t step, uint32_t flags, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    SSMD_TRY(step_dispatch(*p, *d, nullptr, step, flags, static_cast<hipStream_t>(stream)));
    swap_buffers(d);
    return SSMD_OK;
}

int ssmd_derived_update(const ssmd_params *p, const ssmd_fields *d, const ssmd_derived *s, uint64_t step,
                        void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!s) return SSMD_EINVAL;
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    const KParams k = make_kparams(*p);
    Derived v{};
    v.avg[0] = s->avg_phi;
    v.avg[1] = s->avg_ux;
    v.avg[2] = s->avg_uy;
    v.avg[3] = s->avg_uz;
    v.rey[0] = s->avg_uxux;
    v.rey[1] = s->avg_uyuy;
    v.rey[2] = s->avg_uzuz;
    v.rey[3] = s->avg_uxuy;
    v.rey[4] = s->avg_uxuz;
    v.rey[5] = s->avg_uyuz;
    v.vort[0] = s->vort_x;
    v.vort[1] = s->vort_y;
    v.vort[2] = s->vort_z;
    v.vort[3] = s->vort_mag;
    hipLaunchKernelGGL(k_derived, dim3((k.N + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), k, d->phi,
                       d->ux, d->uy, d->uz, v, float(step + 1));
    return herr(hipGetLastError());
}

int ssmd_closure_update(const ssmd_params *p, const ssmd_fields *d, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (p->closure != SSMD_CLOSURE_LEARNED) return SSMD_EINVAL;
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    return launch_closure_update(d->mlp, d->feat, d->tau_t, p->nx * p->ny * p->nz, float(p->tau_scale),
                                 float(p->tau_floor), static_cast<hipStream_t>(stream));
}

int ssmd_launch_phase(const ssmd_params *p, const ssmd_fields *d, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    return launch_phase(*p, *d, static_cast<hipStream_t>(stream));
}

int ssmd_launch_lattice(const ssmd_params *p, const ssmd_fields *d, uint32_t flags, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    return launch_lattice(*p, *d, flags, static_cast<hipStream_t>(stream));
}

int ssmd_launch_boundaries(const ssmd_params *p, const ssmd_fields *d, uint64_t step, int32_t use_counter,
                           uint32_t flags, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    if (use_counter && !d->step_counter) return SSMD_ENOTREADY;
    return launch_boundaries(*p, *d, use_counter ? d->step_counter : nullptr, step, flags,
                             static_cast<hipStream_t>(stream));
}

int ssmd_swap(ssmd_fields *d) {
    if (!d) return SSMD_EINVAL;
    swap_buffers(d);
    return SSMD_OK;
}

int ssmd_run(const ssmd_params *p, ssmd_fields *d, uint64_t step0, uint64_t nsteps, void *stream) {
    for (uint64_t i = 0; i < nsteps; ++i) SSMD_TRY(ssmd_step(p, d, step0 + i, 0, stream));
    return SSMD_OK;
}

int ssmd_graph_create(const ssmd_params *p, const ssmd_fields *d, void *stream, ssmd_graph **out) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d) || !out || !d->step_counter) return SSMD_ENOTREADY;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ssmd_graph *g = new (std::nothrow) ssmd_graph();
    if (!g) return SSMD_EINVAL;
    g->p = *p;
    g->fa = d->f;
    g->fb = d->f_next;
    g->ga = d->g;
    g->gb = d->g_next;
    g->counter = d->step_counter;
    // The LEARNED closure's network pass reads the inputs the lattice kernel
    // wrote and writes tau_t, which the next lattice launch reads; nothing in
    // the boundary kernels or the next phase update touches either.  In the
    // graphs it therefore runs as a branch beside them: forked after the
    // lattice launch, joined before the next one (and at the graph's end).
    const bool fork = p->closure == SSMD_CLOSURE_LEARNED;
    if (fork) {
        int rc = herr(hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking));
        if (rc == 0) rc = herr(hipEventCreateWithFlags(&g->lattice_done, hipEventDisableTiming));
        if (rc == 0) rc = herr(hipEventCreateWithFlags(&g->network_done, hipEventDisableTiming));
        if (rc != 0) {
            ssmd_graph_destroy(g);
            return rc;
        }
    }
    auto forked_step = [&](const ssmd_fields &cur, bool join_first) {
        SSMD_TRY(launch_phase(*p, cur, s, g->counter));
        if (join_first) SSMD_HIP(hipStreamWaitEvent(s, g->network_done, 0));
        SSMD_TRY(launch_lattice(*p, cur, 0, s));
        SSMD_HIP(hipEventRecord(g->lattice_done, s));
        SSMD_HIP(hipStreamWaitEvent(g->side, g->lattice_done, 0));
        SSMD_TRY(launch_closure_update(cur.mlp, cur.feat, cur.tau_t, p->nx * p->ny * p->nz, float(p->tau_scale),
                                       float(p->tau_floor), g->side));
        SSMD_HIP(hipEventRecord(g->network_done, g->side));
        return launch_boundaries(*p, cur, g->counter, 0, 0, s, true);
    };
    // capture `n` steps from the buffers in `cur` (left as after those steps)
    auto capture = [&](ssmd_fields &cur, int n, hipGraph_t *graph, hipGraphExec_t *exec) {
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return int(e);
        int rc = 0;
        bool forked = false;
        for (int i = 0; i < n && rc == 0; ++i) {
            if (fork) {
                rc = forked_step(cur, forked);
                forked = true;
            } else {
                rc = step_dispatch(*p, cur, g->counter, 0, 0, s, true);
            }
            swap_buffers(&cur);
        }
        if (forked) {  // join the last network pass (also after an error: capture must end joined)
            const int rj = herr(hipStreamWaitEvent(s, g->network_done, 0));
            if (rc == 0) rc = rj;
        }
        e = hipStreamEndCapture(s, graph);
        if (rc != 0 || e != hipSuccess) return rc != 0 ? rc : int(e);
        return int(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
    };
    ssmd_fields cur = *d;
    for (int par = 0; par < 2; ++par) {
        ssmd_fields one = cur, block = cur;
        int rc = capture(one, 1, &g->graph[par], &g->exec[par]);
        if (rc == 0) rc = capture(block, kGraphBlock, &g->block_graph[par], &g->block_exec[par]);
        if (rc != 0) {
            ssmd_graph_destroy(g);
            return rc;
        }
        swap_buffers(&cur);
    }
    *out = g;
    return SSMD_OK;
}

int ssmd_graph_launch(ssmd_graph *g, ssmd_fields *d, uint64_t step0, uint64_t nsteps, void *stream) {
    if (!g || !d) return SSMD_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int par;
    if (d->f == g->fa && d->g == g->ga)
        par = 0;
    else if (d->f == g->fb && d->g == g->gb)
        par = 1;
    else
        return SSMD_EINVAL;
    // each captured step's phase kernel advances the counter before its face
    // kernels read it: start one below step0 (modulo 2^64)
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, s, g->counter, step0 - 1);
    SSMD_HIP(hipGetLastError());
    uint64_t i = 0;
    for (; i + kGraphBlock <= nsteps; i += kGraphBlock)  // ends on the same parity
        SSMD_HIP(hipGraphLaunch(g->block_exec[par], s));
    for (; i < nsteps; ++i) {
        SSMD_HIP(hipGraphLaunch(g->exec[par], s));
        swap_buffers(d);
        par ^= 1;
    }
    return SSMD_OK;
}

int ssmd_graph_destroy(ssmd_graph *g) {
    if (!g) return SSMD_OK;
    for (int i = 0; i < 2; ++i) {
        if (g->exec[i]) (void)hipGraphExecDestroy(g->exec[i]);
        if (g->graph[i]) (void)hipGraphDestroy(g->graph[i]);
        if (g->block_exec[i]) (void)hipGraphExecDestroy(g->block_exec[i]);
        if (g->block_graph[i]) (void)hipGraphDestroy(g->block_graph[i]);
    }
    if (g->lattice_done) (void)hipEventDestroy(g->lattice_done);
    if (g->network_done) (void)hipEventDestroy(g->network_done);
    if (g->side) (void)hipStreamDestroy(g->side);
    delete g;
    return SSMD_OK;
}

int ssmd_phase_mass_f64(const ssmd_params *p, const ssmd_fields *d, double *out_host, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d) || !d->reduce) return SSMD_ENOTREADY;
    if (!out_host) return SSMD_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const KParams k = make_kparams(*p);
    const int maxb = (SSMD_REDUCE_DOUBLES - 8) / 4;
    int nb = int((k.N + 255) / 256);
    if (nb > maxb) nb = maxb;
    double *out = d->reduce + SSMD_REDUCE_DOUBLES - 4;
    hipLaunchKernelGGL(k_reduce1, dim3(nb), dim3(256), 0, s, k, d->phi, make_macro(*d), d->reduce);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, s, d->reduce, nb, out);
    SSMD_HIP(hipGetLastError());
    SSMD_HIP(hipMemcpyAsync(out_host, out, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    SSMD_HIP(hipStreamSynchronize(s));
    return SSMD_OK;
}

int ssmd_populations_to_reference(const ssmd_params *p, const ssmd_fields *d, void *f_out, float *g_out,
                                  void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    if ((f_out && (f_out == d->f || f_out == d->f_next)) || (g_out && (g_out == d->g || g_out == d->g_next)))
        return SSMD_EINVAL;  // the conversion is not in place
    if (!f_out && !g_out) return SSMD_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const KParams k = make_kparams(*p);
    const dim3 grid(unsigned((k.N + 255) / 256)), block(256);
    SSMD_TRY(dispatch(*p, [&]<int Q, class P>() {
        using PT = typename P::T;
        hipLaunchKernelGGL((k_export<Q, P>), grid, block, 0, s, k, static_cast<const PT *>(d->f), d->g,
                           static_cast<PT *>(f_out), g_out);
        return herr(hipGetLastError());
    }));
    return SSMD_OK;
}

int ssmd_quantize(int32_t fmt, const float *in, float *out, uint64_t n, void *stream) {
    if (!in || !out) return SSMD_EINVAL;
    if (n == 0) return SSMD_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    switch (fmt) {
        case SSMD_POP_F32: hipLaunchKernelGGL(k_quantize<PopF32>, grid, block, 0, s, in, out, n); break;
        case SSMD_POP_F16: hipLaunchKernelGGL(k_quantize<PopF16>, grid, block, 0, s, in, out, n); break;
        case SSMD_POP_BF16: hipLaunchKernelGGL(k_quantize<PopBF16>, grid, block, 0, s, in, out, n); break;
        default: return SSMD_EINVAL;
    }
    return herr(hipGetLastError());
}


// ---------------------------------------------------------------------------
// Autotuner

int32_t ssmd_auto_lz(const ssmd_params *p, int32_t tile) {
    SSMD_TRY(ssmd_validate(p));
    if (tile < 0 || tile >= kNumTiles) return SSMD_EINVAL;
    return auto_lz(int(p->nx), int(p->ny), int(p->nz), tile);
}

int ssmd_tile_count(void) { return kNumTiles; }

int ssmd_tile_shape(int32_t tile, int32_t *tx, int32_t *ty) {
    if (tile < 0 || tile >= kNumTiles || !tx || !ty) return SSMD_EINVAL;
    *tx = kTiles[tile].tx;
    *ty = kTiles[tile].ty;
    return SSMD_OK;
}

}  // extern "C"

namespace {

// Mean duration (ms) of `reps` launches made by `launch`, after one warm-up.
template <class F>
int time_launches(hipStream_t s, int reps, F &&launch, float *ms) {
    hipEvent_t e0, e1;
    SSMD_HIP(hipEventCreate(&e0));
    SSMD_HIP(hipEventCreate(&e1));
    int rc = launch();
    if (rc == 0) rc = herr(hipEventRecord(e0, s));
    for (int r = 0; r < reps && rc == 0; ++r) rc = launch();
    if (rc == 0) rc = herr(hipEventRecord(e1, s));
    if (rc == 0) rc = herr(hipEventSynchronize(e1));
    float t = 0.f;
    if (rc == 0) rc = herr(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / float(reps > 0 ? reps : 1);
    return rc;
}

// The tuners' trial lattice launches write g_next.  A field set whose g_next
// is g itself (one g buffer) would lose the g the next step reads, so the
// tuners keep a device copy of such buffers and put it back afterwards.
struct GKeeper {
    std::vector<std::pair<float *, size_t>> bufs;  // (g, bytes)
    char *copy = nullptr;
    hipStream_t s = nullptr;
    int save(hipStream_t st) {
        s = st;
        size_t total = 0;
        for (const auto &b : bufs) total += b.second;
        if (total == 0) return SSMD_OK;
        SSMD_HIP(hipMalloc(reinterpret_cast<void **>(&copy), total));
        size_t off = 0;
        for (const auto &b : bufs) {
            SSMD_HIP(hipMemcpyAsync(copy + off, b.first, b.second, hipMemcpyDeviceToDevice, s));
            off += b.second;
        }
        return SSMD_OK;
    }
    int restore() {
        if (!copy) return SSMD_OK;
        int rc = 0;
        size_t off = 0;
        for (const auto &b : bufs) {
            if (rc == 0) rc = herr(hipMemcpyAsync(b.first, copy + off, b.second, hipMemcpyDeviceToDevice, s));
            off += b.second;
        }
        const int rs = herr(hipStreamSynchronize(s));
        (void)hipFree(copy);
        copy = nullptr;
        return rc ? rc : rs;
    }
    ~GKeeper() {
        if (copy) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(copy);
        }
    }
};

}  // namespace

extern "C" {

int ssmd_tune(const ssmd_params *p, const ssmd_fields *d, int32_t reps, void *stream, int32_t *best_tile,
              int32_t *best_lz) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(*p, d)) return SSMD_ENOTREADY;
    if (!best_tile || !best_lz || reps < 1) return SSMD_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    GKeeper keep;
    if (d->g_next == d->g) keep.bufs.push_back({d->g, ssmd_bytes(p, 2)});
    SSMD_TRY(keep.save(s));
    float best = 1e30f;
    int bt = 0, bl = 0;
    const int interior = int(p->nz) - 2;
    const int rc = [&]() -> int {
        SSMD_TRY(launch_phase(*p, *d, s));
        for (int t = 0; t < kNumTiles; ++t) {
            const int lzs[] = {auto_lz(int(p->nx), int(p->ny), int(p->nz), t), 8, 16, 32, 64};
            for (int lz : lzs) {
                if (lz > interior && lz != lzs[0]) continue;
                float ms = 0.f;
                SSMD_TRY(time_launches(s, reps, [&]() { return launch_lattice_sel(*p, *d, 0, t, lz, s); }, &ms));
                if (ms < best) {
                    best = ms;
                    bt = t;
                    bl = lz;
                }
            }
        }
        return SSMD_OK;
    }();
    const int rr = keep.restore();
    SSMD_TRY(rc);
    SSMD_TRY(rr);
    *best_tile = bt;
    *best_lz = bl;
    return SSMD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ensembles

struct ssmd_ensemble {
    int n = 0;
    std::vector<ssmd_params> p;
    std::vector<ssmd_fields> d;     // current descriptors
    std::vector<ssmd_fields> orig;  // descriptors at creation (parity 0)
    int tile = 0;
    int kinds[8] = {};
    int nk = 0, last = -1;
    JetSlot *table[2] = {nullptr, nullptr};
    uint32_t total_tiles = 0, phase_blocks = 0, face_blocks[4] = {0, 0, 0, 0};
    bool phase_quad = false;  // every member's rows a multiple of 4: k_phase_quad_batch
    int parity = 0;
    uint64_t *counter = nullptr;
    uint32_t *queue = nullptr;  // tile queue of the persistent lattice kernel (2 words, zero between launches)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

namespace {

void ensemble_drop_graphs(ssmd_ensemble *e) {
    for (int i = 0; i < 2; ++i) {
        if (e->exec[i]) (void)hipGraphExecDestroy(e->exec[i]);
        if (e->graph[i]) (void)hipGraphDestroy(e->graph[i]);
        e->exec[i] = nullptr;
        e->graph[i] = nullptr;
    }
}

// Host tables for both parities with tile shape `tile`, uploaded to dst[2].
int ensemble_tables(ssmd_ensemble *e, int tile, JetSlot *dst[2], uint32_t *total_tiles) {
    SSMD_HIP(hipDeviceSynchronize());  // no launch may still read the tables being replaced
    std::vector<JetSlot> h[2];
    uint32_t tiles = 0, pblocks = 0, faces[4] = {0, 0, 0, 0};
    const bool mixed = tile == kEnsembleMixedTile;
    const TileShape ts = kTiles[mixed ? 3 : tile];
    bool quad = true;
    for (int j = 0; j < e->n; ++j) quad = quad && e->p[j].nx % 4 == 0;
    uint64_t batch_cells = 0;
    for (int j = 0; j < e->n; ++j) batch_cells += uint64_t(e->p[j].nx) * e->p[j].ny * e->p[j].nz;
    const bool batch_fits = batch_cells * 16u <= (uint64_t(160) << 20);  // as make_kparams' g_cached
    for (int j = 0; j < e->n; ++j) {
        const ssmd_params &pp = e->p[j];
        const ssmd_fields &o = e->orig[j];
        JetSlot js{};
        js.k = make_kparams(pp);
        // the cache policies follow the whole batch: its phase update and
        // lattice launch run over every member's g and phi
        js.k.nt_pull = 0;
        js.k.g_cached = batch_fits ? 1 : 0;
        js.phi = o.phi;
        js.mac = make_macro(o);
        js.narrow = mixed && narrow_member(int(pp.nx)) ? 1 : 0;
        const int tx = js.narrow ? 64 : ts.tx, ty = js.narrow ? 8 : ts.ty;
        js.tiles_x = int((pp.nx + tx - 1) / tx);
        js.tiles_y = int((pp.ny + ty - 1) / ty);
        js.lz = pp.lz > 0 ? pp.lz : auto_lz_shape(int(pp.nx), int(pp.ny), int(pp.nz), tx, ty);
        const int tz = (int(pp.nz) - 2 + js.lz - 1) / js.lz;
        js.tile_begin = tiles;
        tiles += uint32_t(js.tiles_x) * uint32_t(js.tiles_y) * uint32_t(tz);
        js.cell_block_begin = pblocks;
        if (quad)
            pblocks += ((uint32_t(pp.nx / 4) * (pp.ny - 2) + 255u) / 256u) *
                       uint32_t((pp.nz - 2 + kEnsemblePhaseLZ - 1) / kEnsemblePhaseLZ);
        else
            pblocks += ((js.k.sxy + 255u) / 256u) * uint32_t((pp.nz - 2 + kPhaseLZ - 1) / kPhaseLZ);
        for (int ax = 0; ax < 4; ++ax) {
            int na, nb;
            face_extent(ax, js.k.nx, js.k.ny, js.k.nz, na, nb);
            js.face_begin[ax] = faces[ax];
            faces[ax] += uint32_t((na + 63) / 64) * uint32_t((nb + 3) / 4);
        }
        for (int par = 0; par < 2; ++par) {
            JetSlot t = js;
            t.f = par == 0 ? o.f : o.f_next;
            t.f_next = par == 0 ? o.f_next : o.f;
            t.g = par == 0 ? o.g : o.g_next;
            t.g_next = par == 0 ? o.g_next : o.g;
            h[par].push_back(t);
        }
    }
    for (int par = 0; par < 2; ++par)
        SSMD_HIP(hipMemcpy(dst[par], h[par].data(), sizeof(JetSlot) * size_t(e->n), hipMemcpyHostToDevice));
    *total_tiles = tiles;
    e->phase_blocks = pblocks;
    e->phase_quad = quad;
    for (int ax = 0; ax < 4; ++ax) e->face_blocks[ax] = faces[ax];
    return SSMD_OK;
}

int ensemble_launch_step(ssmd_ensemble *e, const JetSlot *tab, int tile, uint32_t total_tiles, uint64_t *counter,
                         uint64_t step, uint32_t flags, hipStream_t s) {
    const ssmd_params &p0 = e->p[0];
    if (e->phase_quad)
        hipLaunchKernelGGL(k_phase_quad_batch, dim3(e->phase_blocks), dim3(256), 0, s, tab, e->n);
    else
        hipLaunchKernelGGL(k_phase_batch, dim3(e->phase_blocks), dim3(256), 0, s, tab, e->n);
    SSMD_HIP(hipGetLastError());
    const bool macros = (flags & SSMD_STEP_MACROS) != 0;
    if (macros && tile != 0) return SSMD_EINVAL;  // tables are laid out for the tile in use
    SSMD_TRY(dispatch(p0, [&]<int Q, class P>() {
        // the mixed tiling runs the 128 x 4 kernel, which reads each member's shape from its slot
        return herr(launch_lattice_batch<Q, P>(tile == kEnsembleMixedTile ? 3 : tile, macros, tab, e->n, total_tiles,
                                               e->queue, s));
    }));
    return dispatch(p0, [&]<int Q, class P>() {
        for (int i = 0; i < e->nk; ++i) {
            const int ax = face_axis_host(e->kinds[i]);
            face_dispatch(e->kinds[i], [&]<int KIND>() {
                const int fl = int(i == e->last) | (macros ? 2 : 0);
                hipLaunchKernelGGL((k_face_batch<Q, P, KIND>), dim3(e->face_blocks[ax], 1, face_layers<Q>(KIND)),
                                   dim3(64, 4), 0, s, tab, e->n,
                                   counter, step, fl);
            });
        }
        if (counter && e->last < 0 && e->nk > 0) hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, s, counter);
        return herr(hipGetLastError());
    });
}

void ensemble_swap(ssmd_ensemble *e) {
    for (auto &f : e->d) swap_buffers(&f);
    e->parity ^= 1;
}

}  // namespace

extern "C" {

int ssmd_ensemble_create(const ssmd_params *params, const ssmd_fields *fields, int32_t n, ssmd_ensemble **out) {
    if (!params || !fields || !out || n < 1 || n > 1024) return SSMD_EINVAL;
    int kinds0[8];
    const int nk0 = face_kinds(params[0], kinds0);
    for (int j = 0; j < n; ++j) {
        SSMD_TRY(ssmd_validate(&params[j]));
        if (!fields_ready(params[j], &fields[j])) return SSMD_ENOTREADY;
        // the batched step has no per-jet network pass
        if (params[j].closure == SSMD_CLOSURE_LEARNED) return SSMD_EUNSUPPORTED;
        if (params[j].q != params[0].q || params[j].pop_format != params[0].pop_format) return SSMD_EUNSUPPORTED;
        int kinds[8];
        const int nk = face_kinds(params[j], kinds);
        if (nk != nk0 || std::memcmp(kinds, kinds0, sizeof(int) * size_t(nk)) != 0) return SSMD_EUNSUPPORTED;
    }
    ssmd_ensemble *e = new (std::nothrow) ssmd_ensemble();
    if (!e) return SSMD_EINVAL;
    e->n = n;
    e->p.assign(params, params + n);
    e->d.assign(fields, fields + n);
    e->orig = e->d;
    e->nk = nk0;
    std::memcpy(e->kinds, kinds0, sizeof(kinds0));
    for (int i = 0; i < nk0; ++i)
        if (!is_inflow(kinds0[i])) e->last = i;
    e->tile = params[0].tile;
    int rc = 0;
    for (int par = 0; par < 2 && rc == 0; ++par)
        rc = herr(hipMalloc(reinterpret_cast<void **>(&e->table[par]), sizeof(JetSlot) * size_t(n)));
    if (rc == 0) rc = herr(hipMalloc(reinterpret_cast<void **>(&e->counter), sizeof(uint64_t)));
    if (rc == 0) rc = herr(hipMalloc(reinterpret_cast<void **>(&e->queue), 2 * sizeof(uint32_t)));
    if (rc == 0) rc = herr(hipMemset(e->queue, 0, 2 * sizeof(uint32_t)));
    if (rc == 0) rc = ensemble_tables(e, e->tile, e->table, &e->total_tiles);
    if (rc != 0) {
        ssmd_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return SSMD_OK;
}

int ssmd_ensemble_set_tile(ssmd_ensemble *e, int32_t tile) {
    if (!e || tile < 0 || tile > kEnsembleMixedTile) return SSMD_EINVAL;
    if (tile == e->tile) return SSMD_OK;
    ensemble_drop_graphs(e);
    e->tile = tile;
    return ensemble_tables(e, tile, e->table, &e->total_tiles);
}

int ssmd_ensemble_step(ssmd_ensemble *e, uint64_t step, uint32_t flags, void *stream) {
    if (!e) return SSMD_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool macros = (flags & SSMD_STEP_MACROS) != 0;
    if (macros && e->tile != 0) {
        // the macro kernel uses tile 0: switch, step, switch back
        const int keep = e->tile;
        SSMD_TRY(ssmd_ensemble_set_tile(e, 0));
        const int rc = ensemble_launch_step(e, e->table[e->parity], 0, e->total_tiles, nullptr, step, flags, s);
        SSMD_TRY(rc);
        ensemble_swap(e);
        SSMD_HIP(hipStreamSynchronize(s));  // the tables are rewritten next
        return ssmd_ensemble_set_tile(e, keep);
    }
    SSMD_TRY(ensemble_launch_step(e, e->table[e->parity], e->tile, e->total_tiles, nullptr, step, flags, s));
    ensemble_swap(e);
    return SSMD_OK;
}

int ssmd_ensemble_graph_launch(ssmd_ensemble *e, uint64_t step0, uint64_t nsteps, void *stream) {
    if (!e) return SSMD_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!e->exec[0]) {
        for (int par = 0; par < 2; ++par) {
            hipError_t err = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (err != hipSuccess) return int(err);
            const int rc = ensemble_launch_step(e, e->table[par], e->tile, e->total_tiles, e->counter, 0, 0, s);
            err = hipStreamEndCapture(s, &e->graph[par]);
            if (rc != 0 || err != hipSuccess) {
                ensemble_drop_graphs(e);
                return rc != 0 ? rc : int(err);
            }
            err = hipGraphInstantiate(&e->exec[par], e->graph[par], nullptr, nullptr, 0);
            if (err != hipSuccess) {
                ensemble_drop_graphs(e);
                return int(err);
            }
        }
    }
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, s, e->counter, step0);
    SSMD_HIP(hipGetLastError());
    for (uint64_t i = 0; i < nsteps; ++i) {
        SSMD_HIP(hipGraphLaunch(e->exec[e->parity], s));
        ensemble_swap(e);
    }
    return SSMD_OK;
}

int ssmd_ensemble_tune(ssmd_ensemble *e, int32_t reps, void *stream, int32_t *best_tile) {
    if (!e || !best_tile || reps < 1) return SSMD_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    GKeeper keep;
    for (int j = 0; j < e->n; ++j)
        if (e->d[j].g_next == e->d[j].g) keep.bufs.push_back({e->d[j].g, ssmd_bytes(&e->p[j], 2)});
    SSMD_TRY(keep.save(s));
    JetSlot *tmp[2] = {nullptr, nullptr};
    int rc = 0;
    for (int par = 0; par < 2 && rc == 0; ++par)
        rc = herr(hipMalloc(reinterpret_cast<void **>(&tmp[par]), sizeof(JetSlot) * size_t(e->n)));
    float best = 1e30f;
    int bt = e->tile;
    const ssmd_params &p0 = e->p[0];
    if (rc == 0) {
        if (e->phase_quad)
            hipLaunchKernelGGL(k_phase_quad_batch, dim3(e->phase_blocks), dim3(256), 0, s, e->table[e->parity], e->n);
        else
            hipLaunchKernelGGL(k_phase_batch, dim3(e->phase_blocks), dim3(256), 0, s, e->table[e->parity], e->n);
        rc = herr(hipGetLastError());
    }
    for (int t = 0; t <= kEnsembleMixedTile && rc == 0; ++t) {
        uint32_t total = 0;
        rc = ensemble_tables(e, t, tmp, &total);
        if (rc != 0) break;
        float ms = 0.f;
        rc = time_launches(s, reps, [&]() {
            return dispatch(p0, [&]<int Q, class P>() {
                return herr(launch_lattice_batch<Q, P>(t == kEnsembleMixedTile ? 3 : t, false, tmp[e->parity], e->n,
                                                       total, e->queue, s));
            });
        }, &ms);
        if (rc == 0 && ms < best) {
            best = ms;
            bt = t;
        }
    }
    for (int par = 0; par < 2; ++par)
        if (tmp[par]) (void)hipFree(tmp[par]);
    const int rr = keep.restore();
    if (rc != 0) return rc;
    SSMD_TRY(rr);
    // restore the face/cell block counts of the tile in use, then switch
    uint32_t total = 0;
    SSMD_TRY(ensemble_tables(e, e->tile, e->table, &total));
    e->total_tiles = total;
    SSMD_TRY(ssmd_ensemble_set_tile(e, bt));
    *best_tile = bt;
    return SSMD_OK;
}

int ssmd_ensemble_fields(const ssmd_ensemble *e, int32_t member, ssmd_fields *out) {
    if (!e || !out || member < 0 || member >= e->n) return SSMD_EINVAL;
    *out = e->d[size_t(member)];
    return SSMD_OK;
}

int ssmd_ensemble_destroy(ssmd_ensemble *e) {
    if (!e) return SSMD_OK;
    ensemble_drop_graphs(e);
    for (int par = 0; par < 2; ++par)
        if (e->table[par]) (void)hipFree(e->table[par]);
    if (e->counter) (void)hipFree(e->counter);
    if (e->queue) (void)hipFree(e->queue);
    delete e;
    return SSMD_OK;
}

}  // extern "C"
