// Test driver for the CPU oracle under AddressSanitizer and
// UndefinedBehaviorSanitizer (built and run by tests/test_oracle_sanitize.py;
// host code only).  Each argument names a file holding one raw ssmd_params;
// for each, the oracle is stepped in FP64 and in FP32 through every extern
// "C" entry point (whole steps, the kernels one by one, field and population
// access), then the closure network's forward, gradient and Adam step run
// once.  A memory error or undefined behaviour aborts the process.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "ssmd.h"

extern "C" {
struct oracle_handle;
size_t oracle_params_size(void);
oracle_handle *oracle_create(const ssmd_params *p, int fp64);
void oracle_destroy(oracle_handle *h);
void oracle_set_threads(oracle_handle *h, int n);
void oracle_init(oracle_handle *h);
void oracle_run(oracle_handle *h, uint64_t step0, uint64_t nsteps);
void oracle_k_phase(oracle_handle *h);
void oracle_k_forces(oracle_handle *h);
void oracle_k_moments(oracle_handle *h);
void oracle_k_collide(oracle_handle *h);
void oracle_k_boundaries(oracle_handle *h, uint64_t step);
int oracle_get_field(oracle_handle *h, int id, double *out);
int oracle_set_field(oracle_handle *h, int id, const double *in);
int oracle_get_pops(oracle_handle *h, int which, double *out);
int oracle_set_pops(oracle_handle *h, int which, const double *in);
void oracle_sums(oracle_handle *h, double *out4);
int oracle_mlp_params(void);
void oracle_mlp_forward(const double *params, const double *x, int64_t n, int quant, double *y);
void oracle_mlp_grad(const double *params, const double *x, const double *t, const double *w, int64_t n, int quant,
                     double *grad, double *loss);
void oracle_adam(double *params, const double *grad, double *m, double *v, int64_t n, int64_t step, double lr,
                 double b1, double b2, double eps);
}

static bool run_case(const ssmd_params &p, int fp64, int steps) {
    oracle_handle *h = oracle_create(&p, fp64);
    if (!h) return false;
    oracle_set_threads(h, 2);
    oracle_init(h);
    oracle_run(h, 0, uint64_t(steps));
    // one more step kernel by kernel, through the field and population accessors
    const size_t n = size_t(p.nx) * p.ny * p.nz;
    std::vector<double> field(n), pops(size_t(p.q) * n), g(7 * n);
    bool ok = oracle_get_field(h, 10, field.data()) == 0 && oracle_set_field(h, 10, field.data()) == 0 &&
              oracle_get_pops(h, 0, pops.data()) == 0 && oracle_set_pops(h, 0, pops.data()) == 0 &&
              oracle_get_pops(h, 1, g.data()) == 0 && oracle_set_pops(h, 1, g.data()) == 0;
    oracle_k_phase(h);
    oracle_k_forces(h);
    oracle_k_moments(h);
    oracle_k_collide(h);
    oracle_k_boundaries(h, uint64_t(steps));
    double s[4];
    oracle_sums(h, s);
    std::printf("case %ux%ux%u q%d fmt %d closure %d fp64 %d: sums %.9g %.9g %.9g %.9g\n", p.nx, p.ny, p.nz, p.q,
                p.pop_format, p.closure, fp64, s[0], s[1], s[2], s[3]);
    for (double v : s) ok = ok && std::isfinite(v);
    oracle_destroy(h);
    return ok;
}

int main(int argc, char **argv) {
    if (oracle_params_size() != sizeof(ssmd_params)) return 3;
    bool ok = true;
    for (int a = 1; a < argc; ++a) {
        ssmd_params p;
        FILE *f = std::fopen(argv[a], "rb");
        if (!f) return 4;
        const size_t got = std::fread(&p, 1, sizeof p, f);
        std::fclose(f);
        if (got != sizeof p) return 5;
        for (int fp64 = 1; fp64 >= 0; --fp64) ok = run_case(p, fp64, 6) && ok;
    }
    // closure network: forward, gradient and one Adam step on 40 samples
    const int np = oracle_mlp_params();
    std::vector<double> w(np), grad(np), m(np, 0.0), v(np, 0.0), x(40 * 8), t(40), y(40);
    for (int i = 0; i < np; ++i) w[i] = 0.05 * std::sin(0.7 * i);
    for (int i = 0; i < 40 * 8; ++i) x[i] = std::cos(0.3 * i);
    for (int i = 0; i < 40; ++i) t[i] = 0.01 * i;
    double loss = 0;
    oracle_mlp_forward(w.data(), x.data(), 40, 1, y.data());
    oracle_mlp_grad(w.data(), x.data(), t.data(), nullptr, 40, 1, grad.data(), &loss);
    oracle_adam(w.data(), grad.data(), m.data(), v.data(), np, 1, 1e-3, 0.9, 0.999, 1e-8);
    std::printf("network: loss %.9g\n", loss);
    ok = ok && std::isfinite(loss);
    std::printf(ok ? "sanitized oracle run: ok\n" : "sanitized oracle run: FAILED\n");
    return ok ? 0 : 1;
}
