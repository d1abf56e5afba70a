// ssmd_run -- native host driver of the gfx950 solver (the reference's main.cu
// flow: main.cu:35-213), calling the kernels only through the C ABI.
//
//   ssmd_run <D3Q19|D3Q27> <ID> [options]
//     --case reference|jet_small|jet_d3q27|jet_large   parameter block (default: reference)
//     --steps N          time steps after t=0 (default 1000); runs N+1 steps as the reference loop
//     --save-every M     .bin/.vti output interval (default 1000; 0 = none, the BENCHMARK build)
//     --pop f32|f16|bf16 population storage (default f32)
//     --out DIR          output root (default ./bin); files go to DIR/<VS>/<ID>/
//     --no-tune          keep the default launch shape (else ssmd_tune picks the lattice
//                        tile and z-chunk on the t = 0 state before the graphs are made)
//   GPU_INDEX selects the device (hostFunctions.cuh:139-172).
//
// Prints the metadata block, "Step N" at saves, and the final time and MLUPS
// (all cells, NSTEPS+1 steps; main.cu:199-208).  Aborts with a message on any
// nonzero return from the library, as the reference's checkCudaErrors does.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

#include "ssmd.h"

namespace {

void check(int rc, const char *what) {
    if (rc != 0) {
        std::fprintf(stderr, "ssmd error at %s: [%d] %s\n", what, rc, ssmd_error_string(rc));
        std::fflush(stderr);
        std::abort();
    }
}

void usage(const char *argv0) {
    std::cerr << "Usage: " << argv0
              << " <D3Q19|D3Q27> <ID> [--case reference|jet_small|jet_d3q27|jet_large] [--steps N]"
                 " [--save-every M] [--pop f32|f16|bf16] [--out DIR] [--no-tune]\n";
}

// Single bottom-face jet: tau, tau_g, width, sigma, diameter, u, seed.
void jet(ssmd_params &p, uint32_t nx, uint32_t ny, uint32_t nz, int q, double sigma, int diameter, double u,
         uint64_t seed) {
    std::memset(&p, 0, sizeof(p));
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.q = q;
    const double tau = 0.55, tau_g = 0.8, width = 4.0;
    p.visc_water = p.visc_oil = (tau - 0.5) / 3.0;
    p.gamma = 4.0 * 0.25 * (tau_g - 0.5) / width / 0.25;
    p.sigma = sigma;
    p.perturb = 0.01;
    p.seed = seed;
    p.oil.enabled = 1;
    p.oil.center_a = (nx - 1) * 0.5;
    p.oil.center_b = (ny - 1) * 0.5;
    p.oil.r2 = double((diameter / 2) * (diameter / 2));
    p.oil.diameter = diameter;
    p.oil.u = u;
    p.oil.phi = 1.0;
    p.outflow_z = 1;
    p.wrap_x = p.wrap_y = SSMD_WRAP_REFERENCE;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        usage(argv[0]);
        return 1;
    }
    const std::string vs = argv[1], id = argv[2];
    if (vs != "D3Q19" && vs != "D3Q27") {
        usage(argv[0]);
        return 1;
    }
    const int q = vs == "D3Q19" ? 19 : 27;
    std::string kase = "reference", pop = "f32", out_root = "bin";
    uint64_t nsteps = 1000, save_every = 1000;
    bool tune = true;
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage(argv[0]);
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--case") kase = next();
        else if (a == "--steps") nsteps = std::stoull(next());
        else if (a == "--save-every") save_every = std::stoull(next());
        else if (a == "--pop") pop = next();
        else if (a == "--out") out_root = next();
        else if (a == "--no-tune") tune = false;
        else {
            usage(argv[0]);
            return 1;
        }
    }

    ssmd_params p;
    if (kase == "reference") check(ssmd_params_reference_default(&p, q), "params");
    else if (kase == "jet_small") jet(p, 64, 64, 128, q, 0.1, 8, 0.05, 1);
    else if (kase == "jet_d3q27") jet(p, 128, 128, 512, q, 0.05, 16, 0.05, 2);
    else if (kase == "jet_large") jet(p, 256, 256, 1024, q, 0.05, 24, 0.04, 3);
    else {
        usage(argv[0]);
        return 1;
    }
    p.pop_format = pop == "f16" ? SSMD_POP_F16 : (pop == "bf16" ? SSMD_POP_BF16 : SSMD_POP_F32);
    check(ssmd_validate(&p), "validate");

    // createSimulationDirectory (hostFunctions.cuh:33-45)
    const std::filesystem::path dir = std::filesystem::path(out_root) / vs / id;
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);

    // setDeviceFromEnv (hostFunctions.cuh:139-172)
    int dev = 0;
    if (const char *env = std::getenv("GPU_INDEX")) {
        char *end = nullptr;
        const long v = std::strtol(env, &end, 10);
        if (end != env && v >= 0) dev = int(v);
    }
    if (hipSetDevice(dev) != hipSuccess) {
        std::fprintf(stderr, "hipSetDevice(%d) failed\n", dev);
        return 1;
    }
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) std::printf("Using GPU %d: %s (%s)\n", dev, prop.name, prop.gcnArchName);

    ssmd_fields d;
    check(ssmd_alloc(&p, &d), "alloc");
    hipStream_t s{};
    check(int(hipStreamCreate(&s)), "stream");
    check(ssmd_init(&p, &d, s), "init");
    check(int(hipStreamSynchronize(s)), "init sync");

    check(ssmd_write_info(&p, dir.c_str(), id.c_str(), vs.c_str(), nsteps, save_every), "info");
    std::cout << "\n---------------------------- SIMULATION METADATA ----------------------------\n"
              << "Velocity set:       " << vs << '\n'
              << "Case:               " << kase << '\n'
              << "Populations:        " << pop << '\n'
              << "NX:                 " << p.nx << '\n'
              << "NY:                 " << p.ny << '\n'
              << "NZ:                 " << p.nz << '\n'
              << "Cells:              " << uint64_t(p.nx) * p.ny * p.nz << '\n'
              << "-----------------------------------------------------------------------------\n\n";

    if (tune) {  // launch shape only: the state the first step reads is unchanged
        int32_t tile = 0, lz = 0;
        check(ssmd_tune(&p, &d, 10, s, &tile, &lz), "tune");
        p.tile = tile;
        p.lz = lz;
        std::cout << "Lattice launch shape: tile " << tile << ", z-chunk " << lz << " (ssmd_tune)\n";
    }
    ssmd_graph *g = nullptr;
    check(ssmd_graph_create(&p, &d, s, &g), "graph");
    const char *names[] = {"phi", "rho"};  // main.cu:126-129

    // NaN/Inf sentinel (the reference has none): the FP64 sums of phi and
    // rho u after a step that wrote every moment must be finite
    auto sentinel = [&](uint64_t step) {
        double sums[4];
        check(ssmd_phase_mass_f64(&p, &d, sums, s), "sentinel");
        for (double v : sums)
            if (!std::isfinite(v)) {
                std::cerr << "non-finite state at step " << step << ": sum phi " << sums[0] << ", sum rho u ("
                          << sums[1] << ", " << sums[2] << ", " << sums[3] << ")\n";
                std::exit(2);
            }
    };

    auto save = [&](uint64_t step) {
        // a step that writes every moment field, then the sentinel and the files
        check(ssmd_step(&p, &d, step, SSMD_STEP_MACROS, s), "macro step");
        check(int(hipStreamSynchronize(s)), "sync");
        sentinel(step);
        std::cout << "Step " << step << ": bins in " << dir.string() << "\n";
        check(ssmd_write_bin(&p, d.phi, dir.c_str(), "phi", step, s), "bin phi");
        check(ssmd_write_bin(&p, d.rho, dir.c_str(), "rho", step, s), "bin rho");
        check(ssmd_write_vti(&p, dir.c_str(), step, names, 2), "vti");
    };

    const auto t0 = std::chrono::high_resolution_clock::now();
    uint64_t step = 0;
    while (step <= nsteps) {
        const bool is_save = save_every && (step % save_every == 0 || step == nsteps);
        if (is_save) {
            save(step);
            ++step;
            continue;
        }
        // replay graphs up to the next save step (or the end)
        uint64_t stop = nsteps + 1;
        if (save_every) {
            const uint64_t nxt = (step / save_every + 1) * save_every;
            if (nxt < stop) stop = nxt;
            if (nsteps < stop) stop = nsteps;
        }
        check(ssmd_graph_launch(g, &d, step, stop - step, s), "graph launch");
        step = stop;
    }
    check(int(hipStreamSynchronize(s)), "final sync");
    const auto t1 = std::chrono::high_resolution_clock::now();
    const double elapsed = std::chrono::duration<double>(t1 - t0).count();
    const double lups = double(p.nx) * p.ny * p.nz * double(nsteps + 1);
    std::cout << "\n// =============================================== //\n";
    std::cout << "     Total execution time    : " << elapsed << " s\n";
    std::cout << "     Performance             : " << lups / (elapsed * 1e6) << " MLUPS\n";
    std::cout << "// =============================================== //\n\n";
    sentinel(nsteps);
    double mass[4];
    check(ssmd_phase_mass_f64(&p, &d, mass, s), "mass");
    std::cout << "     Phase mass (FP64)       : " << mass[0] << "\n";
    check(ssmd_graph_destroy(g), "graph destroy");
    check(int(hipStreamDestroy(s)), "stream destroy");
    check(ssmd_free(&d), "free");
    return 0;
}
