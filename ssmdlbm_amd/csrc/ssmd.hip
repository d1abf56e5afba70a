// ssmd.hip -- C ABI of the gfx950 hot path (declared in include/ssmd.h).
//
// Host responsibilities mirrored from the reference driver:
//   FieldAllocate (fieldAllocate/FieldAllocate.cuh:42-162)  -> ssmd_alloc/free
//   init kernels (main.cu:109-112)                           -> ssmd_init
//   graph + boundary launches per step (main.cu:159-172)     -> ssmd_step
//   graph::captureGraph (cuda/CUDAGraph.cuh:34-57)           -> ssmd_graph_*
// Errors are returned, never aborted on.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/ssmd.h"
#include "kernels.hpp"

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
    // relaxation:: (globalFunctions.cuh:334-385): formed in double, cast.
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
    return m;
}

bool fields_ready(const ssmd_fields *d) {
    return d && d->rho && d->ux && d->uy && d->uz && d->pxx && d->pyy && d->pzz && d->pxy && d->pxz && d->pyz &&
           d->phi && d->fsx && d->fsy && d->fsz && d->f && d->g && d->f_next && d->g_next;
}

// Launch shape of the fused lattice kernel.
constexpr int kTX = 64, kTY = 4;

int lattice_lz(const ssmd_params &p) {
    // enough z-chunks that the grid holds >= ~2048 blocks on large boxes,
    // without letting the 4-plane halo prologue dominate a chunk
    const int interior = int(p.nz) - 2;
    const long cols = long((p.nx + kTX - 1) / kTX) * long((p.ny + kTY - 1) / kTY);
    int lz = interior;
    while (lz > 16 && cols * ((interior + lz - 1) / lz) < 2048) lz = (lz + 1) / 2;
    return lz < 1 ? 1 : lz;
}

template <int Q, class P, bool MAC>
int launch_lattice_t(const ssmd_params &p, const KParams &k, const ssmd_fields &d, hipStream_t s) {
    const int lz = lattice_lz(p);
    dim3 grid((p.nx + kTX - 1) / kTX, (p.ny + kTY - 1) / kTY, (p.nz - 2 + lz - 1) / lz);
    dim3 block(kTX, kTY, 1);
    hipLaunchKernelGGL((k_lattice<Q, P, kTX, kTY, MAC>), grid, block, 0, s, k,
                       static_cast<const typename P::T *>(d.f), static_cast<typename P::T *>(d.f_next), d.g_next,
                       d.phi, make_macro(d), lz);
    return herr(hipGetLastError());
}

template <int Q, class P>
int launch_lattice_q(const ssmd_params &p, const KParams &k, const ssmd_fields &d, bool mac, hipStream_t s) {
    return mac ? launch_lattice_t<Q, P, true>(p, k, d, s) : launch_lattice_t<Q, P, false>(p, k, d, s);
}

template <int Q, class P>
int launch_bcs_t(const ssmd_params &p, const KParams &k, const ssmd_fields &d, uint64_t *counter, uint64_t step,
                 hipStream_t s) {
    using PT = typename P::T;
    PT *fo = static_cast<PT *>(d.f_next);
    const Macro m = make_macro(d);
    const dim3 b2(64, 4, 1);
    auto g2 = [&](uint32_t na, uint32_t nb) { return dim3((na + 63) / 64, (nb + 3) / 4, 1); };
    // which kernel is last (it advances the device step counter)
    int last = 0;  // 1 outflowY, 2 outflowZ, 3 px, 4 py, 5 pz
    if (p.outflow_y) last = 1;
    if (p.outflow_z) last = 2;
    if (p.wrap_x) last = 3;
    if (p.wrap_y) last = 4;
    if (p.wrap_z) last = 5;
    if (p.water.enabled)
        hipLaunchKernelGGL((k_inflow<Q, P, 1>), g2(p.nx, p.nz), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, step);
    if (p.oil.enabled)
        hipLaunchKernelGGL((k_inflow<Q, P, 2>), g2(p.nx, p.ny), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, step);
    if (p.outflow_y)
        hipLaunchKernelGGL((k_outflow<Q, P, 1>), g2(p.nx, p.nz), b2, 0, s, k, fo, d.g_next, d.phi, m, counter,
                           int(last == 1));
    if (p.outflow_z)
        hipLaunchKernelGGL((k_outflow<Q, P, 2>), g2(p.nx, p.ny), b2, 0, s, k, fo, d.g_next, d.phi, m, counter,
                           int(last == 2));
    auto per = [&](int axis, int mode, uint32_t na, uint32_t nb, int me) {
        const int bl = int(last == me);
        if (axis == 0 && mode == SSMD_WRAP_REFERENCE)
            hipLaunchKernelGGL((k_periodic<Q, P, 0, 1>), g2(na, nb), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, bl);
        if (axis == 0 && mode == SSMD_WRAP_TRUE)
            hipLaunchKernelGGL((k_periodic<Q, P, 0, 2>), g2(na, nb), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, bl);
        if (axis == 1 && mode == SSMD_WRAP_REFERENCE)
            hipLaunchKernelGGL((k_periodic<Q, P, 1, 1>), g2(na, nb), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, bl);
        if (axis == 1 && mode == SSMD_WRAP_TRUE)
            hipLaunchKernelGGL((k_periodic<Q, P, 1, 2>), g2(na, nb), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, bl);
        if (axis == 2 && mode == SSMD_WRAP_REFERENCE)
            hipLaunchKernelGGL((k_periodic<Q, P, 2, 1>), g2(na, nb), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, bl);
        if (axis == 2 && mode == SSMD_WRAP_TRUE)
            hipLaunchKernelGGL((k_periodic<Q, P, 2, 2>), g2(na, nb), b2, 0, s, k, fo, d.g_next, d.phi, m, counter, bl);
    };
    if (p.wrap_x) per(0, p.wrap_x, p.ny, p.nz, 3);
    if (p.wrap_y) per(1, p.wrap_y, p.nx, p.nz, 4);
    if (p.wrap_z) per(2, p.wrap_z, p.nx, p.ny, 5);
    if (counter && last == 0 && (p.water.enabled || p.oil.enabled)) hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, s, counter);
    return herr(hipGetLastError());
}

int launch_phase(const ssmd_params &p, const ssmd_fields &d, hipStream_t s) {
    const KParams k = make_kparams(p);
    hipLaunchKernelGGL(k_phase, dim3((k.N + 255) / 256), dim3(256), 0, s, k, d.g, d.phi);
    return herr(hipGetLastError());
}

// Calls fn.template operator()<Q, P>() for the parameter block's lattice and
// storage format.
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

int launch_lattice(const ssmd_params &p, const ssmd_fields &d, uint32_t flags, hipStream_t s) {
    const KParams k = make_kparams(p);
    return dispatch(p, [&]<int Q, class P>() { return launch_lattice_q<Q, P>(p, k, d, (flags & SSMD_STEP_MACROS) != 0, s); });
}

int launch_boundaries(const ssmd_params &p, const ssmd_fields &d, uint64_t *counter, uint64_t step, hipStream_t s) {
    const KParams k = make_kparams(p);
    return dispatch(p, [&]<int Q, class P>() { return launch_bcs_t<Q, P>(p, k, d, counter, step, s); });
}

// One step on the buffers as they are in *d (no swap).
int step_dispatch(const ssmd_params &p, const ssmd_fields &d, uint64_t *counter, uint64_t step, uint32_t flags,
                  hipStream_t s) {
    SSMD_TRY(launch_phase(p, d, s));
    SSMD_TRY(launch_lattice(p, d, flags, s));
    return launch_boundaries(p, d, counter, step, s);
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

struct ssmd_graph {
    ssmd_params p;
    void *fa, *fb;  // buffer identities: parity 0 reads fa
    float *ga, *gb;
    hipGraphExec_t exec[2];
    hipGraph_t graph[2];
    uint64_t *counter;
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
    p->water.u = u_water;
    p->water.phi = 0.0;
    p->oil.enabled = 1;
    p->oil.center_a = cx;
    p->oil.center_b = double(0.5f * float(cy));
    p->oil.r2 = double(radius * radius);
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
    const uint64_t n = uint64_t(p->nx) * p->ny * p->nz;
    if (n >= (uint64_t(1) << 32) - uint64_t(p->nx) * p->ny) return SSMD_EUNSUPPORTED;
    for (int w : {p->wrap_x, p->wrap_y, p->wrap_z})
        if (w < SSMD_WRAP_NONE || w > SSMD_WRAP_TRUE) return SSMD_EINVAL;
    if (p->wrap_y && (p->water.enabled || p->outflow_y)) return SSMD_EINVAL;
    if (p->wrap_z && (p->oil.enabled || p->outflow_z)) return SSMD_EINVAL;
    if (!(p->visc_water > 0.0) || !(p->visc_oil > 0.0)) return SSMD_EINVAL;
    return SSMD_OK;
}

size_t ssmd_bytes(const ssmd_params *p, int32_t which) {
    if (ssmd_validate(p) != 0) return 0;
    const size_t n = size_t(p->nx) * p->ny * p->nz;
    if (which == 0) return n * 4;
    if (which == 1) return n * size_t(p->q) * pop_size(p->pop_format);
    if (which == 2) return n * 7 * 4;
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
    float **gs[] = {&d->g, &d->g_next};
    for (float **q : gs) {
        hipError_t e = hipMalloc(reinterpret_cast<void **>(q), gb);
        if (e != hipSuccess) return fail(e);
        e = hipMemset(*q, 0, gb);
        if (e != hipSuccess) return fail(e);
    }
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&d->step_counter), sizeof(uint64_t));
    if (e != hipSuccess) return fail(e);
    e = hipMemset(d->step_counter, 0, sizeof(uint64_t));
    if (e != hipSuccess) return fail(e);
    e = hipMalloc(reinterpret_cast<void **>(&d->reduce), SSMD_REDUCE_DOUBLES * sizeof(double));
    if (e != hipSuccess) return fail(e);
    return SSMD_OK;
}

int ssmd_free(ssmd_fields *d) {
    if (!d) return SSMD_EINVAL;
    if (!d->owned) return SSMD_OK;
    void *ptrs[] = {d->rho, d->ux,  d->uy,  d->uz,  d->pxx, d->pyy, d->pzz,    d->pxy,         d->pxz,
                    d->pyz, d->phi, d->fsx, d->fsy, d->fsz, d->f,   d->f_next, d->g,           d->g_next,
                    d->step_counter, d->reduce};
    int rc = 0;
    for (void *q : ptrs)
        if (q) {
            const int r = herr(hipFree(q));
            if (r && !rc) rc = r;
        }
    std::memset(d, 0, sizeof(*d));
    return rc;
}

int ssmd_init(const ssmd_params *p, ssmd_fields *d, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(d)) return SSMD_ENOTREADY;
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
    if (!fields_ready(d)) return SSMD_ENOTREADY;
    SSMD_TRY(step_dispatch(*p, *d, nullptr, step, flags, static_cast<hipStream_t>(stream)));
    swap_buffers(d);
    return SSMD_OK;
}

int ssmd_launch_phase(const ssmd_params *p, const ssmd_fields *d, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(d)) return SSMD_ENOTREADY;
    return launch_phase(*p, *d, static_cast<hipStream_t>(stream));
}

int ssmd_launch_lattice(const ssmd_params *p, const ssmd_fields *d, uint32_t flags, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(d)) return SSMD_ENOTREADY;
    return launch_lattice(*p, *d, flags, static_cast<hipStream_t>(stream));
}

int ssmd_launch_boundaries(const ssmd_params *p, const ssmd_fields *d, uint64_t step, int32_t use_counter,
                           void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(d)) return SSMD_ENOTREADY;
    if (use_counter && !d->step_counter) return SSMD_ENOTREADY;
    return launch_boundaries(*p, *d, use_counter ? d->step_counter : nullptr, step, static_cast<hipStream_t>(stream));
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
    if (!fields_ready(d) || !out || !d->step_counter) return SSMD_ENOTREADY;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ssmd_graph *g = new (std::nothrow) ssmd_graph();
    if (!g) return SSMD_EINVAL;
    g->p = *p;
    g->fa = d->f;
    g->fb = d->f_next;
    g->ga = d->g;
    g->gb = d->g_next;
    g->counter = d->step_counter;
    ssmd_fields cur = *d;
    for (int par = 0; par < 2; ++par) {
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            ssmd_graph_destroy(g);
            return int(e);
        }
        const int rc = step_dispatch(*p, cur, g->counter, 0, 0, s);
        e = hipStreamEndCapture(s, &g->graph[par]);
        if (rc != 0 || e != hipSuccess) {
            ssmd_graph_destroy(g);
            return rc != 0 ? rc : int(e);
        }
        e = hipGraphInstantiate(&g->exec[par], g->graph[par], nullptr, nullptr, 0);
        if (e != hipSuccess) {
            ssmd_graph_destroy(g);
            return int(e);
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
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, s, g->counter, step0);
    SSMD_HIP(hipGetLastError());
    for (uint64_t i = 0; i < nsteps; ++i) {
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
    }
    delete g;
    return SSMD_OK;
}

int ssmd_phase_mass_f64(const ssmd_params *p, const ssmd_fields *d, double *out_host, void *stream) {
    SSMD_TRY(ssmd_validate(p));
    if (!fields_ready(d) || !d->reduce) return SSMD_ENOTREADY;
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

}  // extern "C"
