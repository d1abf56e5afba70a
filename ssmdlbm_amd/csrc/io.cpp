// io.cpp -- host-side output of the solver fields, same files as the
// reference's post-processing:
//   ssmd_write_bin   host::PostProcess::bin  (postProcess/PostProcess.cuh:45-98):
//                    one raw float32 file <dir>/<name><step:06>.bin per field
//   ssmd_write_vti   host::PostProcess::vti  (PostProcess.cuh:101-262): VTK
//                    ImageData step_<step:06>.vti, appended raw encoding, a
//                    uint32 byte count before each array, re-reading the .bin files
//   ssmd_write_info  host::generateSimulationInfoFile (functions/hostFunctions.cuh:47-87)
//   ssmd_checkpoint_write / _read: restart files (no reference counterpart)
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ssmd.h"

namespace {

std::string step6(uint64_t step) {
    std::ostringstream ss;
    ss << std::setw(6) << std::setfill('0') << step;
    return ss.str();
}

// Restart file: header, then the buffers below in order, raw (for the
// LEARNED closure the last one is the packed network).
constexpr char kMagic[8] = {'S', 'S', 'M', 'D', 'C', 'K', 'P', '2'};
struct CkptHeader {
    char magic[8];
    int32_t abi, params_bytes;
    uint64_t step;
    ssmd_params p;
};

struct Buf {
    void *ptr;
    size_t bytes;
};

// f, g as the next step reads them, phi, the 13 moment fields; for the
// LEARNED closure also tau_t and the packed network the next step applies.
std::vector<Buf> state_buffers(const ssmd_params &p, const ssmd_fields &d) {
    const size_t n = size_t(p.nx) * p.ny * p.nz;
    const size_t sf = p.pop_format == SSMD_POP_F32 ? 4 : 2;
    std::vector<Buf> b = {{d.f, n * size_t(p.q) * sf}, {d.g, n * 4 * 4}};
    float *scal[] = {d.phi, d.rho, d.ux, d.uy, d.uz, d.pxx, d.pyy, d.pzz, d.pxy, d.pxz, d.pyz, d.fsx, d.fsy, d.fsz};
    for (float *q : scal) b.push_back({q, n * 4});
    if (p.closure == SSMD_CLOSURE_LEARNED) {
        b.push_back({d.tau_t, n * 4});
        b.push_back({d.mlp, ssmd_mlp_packed_bytes()});
    }
    return b;
}

bool same_jet(const ssmd_jet &a, const ssmd_jet &b) {
    return a.enabled == b.enabled && a.diameter == b.diameter && a.center_a == b.center_a &&
           a.center_b == b.center_b && a.r2 == b.r2 && a.u == b.u && a.phi == b.phi;
}

// Every parameter of the case; only the launch shape (tile, lz) may differ.
bool same_case(const ssmd_params &a, const ssmd_params &b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.q == b.q && a.pop_format == b.pop_format &&
           a.outflow_y == b.outflow_y && a.outflow_z == b.outflow_z && a.wrap_x == b.wrap_x &&
           a.wrap_y == b.wrap_y && a.wrap_z == b.wrap_z && a.visc_water == b.visc_water &&
           a.visc_oil == b.visc_oil && a.sigma == b.sigma && a.gamma == b.gamma && a.perturb == b.perturb &&
           a.seed == b.seed && same_jet(a.water, b.water) && same_jet(a.oil, b.oil) && a.closure == b.closure &&
           a.cs == b.cs && a.strain_scale == b.strain_scale && a.grad_scale == b.grad_scale &&
           a.tau_scale == b.tau_scale && a.tau_floor == b.tau_floor && a.flux_coeff == b.flux_coeff;
}

// Nominal jet diameter: the stated one, else the even diameter of the disc.
double jet_diameter(const ssmd_jet &j) { return j.diameter > 0 ? double(j.diameter) : 2.0 * std::floor(std::sqrt(j.r2) + 1e-9); }

// A dimensionless number as the reference prints its integer constants
// (constants.cuh:86-90): an integer when it is one to 1e-6, else the value.
std::string dimensionless(double v) {
    std::ostringstream ss;
    const double r = std::round(v);
    if (std::isfinite(v) && std::fabs(v - r) <= 1e-6 * std::fabs(v)) ss << static_cast<long long>(r);
    else ss << v;
    return ss.str();
}

constexpr size_t kChunk = size_t(64) << 20;

}  // namespace

extern "C" {

int ssmd_write_bin(const ssmd_params *p, const float *device_field, const char *dir, const char *name, uint64_t step,
                   void *stream) {
    if (!p || !dir || !name) return SSMD_EINVAL;
    if (ssmd_validate(p) != 0) return SSMD_EINVAL;
    if (!device_field) return SSMD_OK;  // a field never allocated: skipped, as PostProcess::bin does
    const size_t n = size_t(p->nx) * p->ny * p->nz;
    std::vector<float> h(n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(h.data(), device_field, n * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return int(e);
    const std::filesystem::path out = std::filesystem::path(dir) / (std::string(name) + step6(step) + ".bin");
    std::ofstream f(out, std::ios::binary | std::ios::trunc);
    if (!f) return SSMD_EINVAL;
    f.write(reinterpret_cast<const char *>(h.data()), std::streamsize(n * sizeof(float)));
    return f ? SSMD_OK : SSMD_EINVAL;
}

int ssmd_write_vti(const ssmd_params *p, const char *dir, uint64_t step, const char *const *names, int32_t count) {
    if (!p || !dir || !names || count <= 0) return SSMD_EINVAL;
    const uint64_t node_bytes = uint64_t(p->nx) * p->ny * p->nz * sizeof(float);
    struct Arr {
        std::string name;
        std::filesystem::path path;
        uint64_t offset;
    };
    std::vector<Arr> arrays;
    uint64_t off = 0;
    for (int i = 0; i < count; ++i) {
        const std::filesystem::path bin = std::filesystem::path(dir) / (std::string(names[i]) + step6(step) + ".bin");
        std::error_code ec;
        if (!std::filesystem::exists(bin, ec) || ec) continue;  // missing .bin: skipped, as the reference
        const uint64_t fs = std::filesystem::file_size(bin, ec);
        if (ec || fs != node_bytes) continue;  // wrong size: skipped
        arrays.push_back({names[i], bin, off});
        off += sizeof(uint32_t) + node_bytes;
    }
    if (arrays.empty()) return SSMD_EINVAL;
    const std::filesystem::path out = std::filesystem::path(dir) / ("step_" + step6(step) + ".vti");
    std::ofstream v(out, std::ios::binary | std::ios::trunc);
    if (!v) return SSMD_EINVAL;
    const double zero = 0.0, one = 1.0;
    v << "<?xml version=\"1.0\"?>\n";
    v << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    v << "  <ImageData WholeExtent=\"0 " << (p->nx - 1) << " 0 " << (p->ny - 1) << " 0 " << (p->nz - 1)
      << "\" Origin=\"" << zero << " " << zero << " " << zero << "\" Spacing=\"" << one << " " << one << " " << one
      << "\">\n";
    v << "    <Piece Extent=\"0 " << (p->nx - 1) << " 0 " << (p->ny - 1) << " 0 " << (p->nz - 1) << "\">\n";
    v << "      <PointData Scalars=\"" << arrays.front().name << "\">\n";
    for (const Arr &a : arrays)
        v << "        <DataArray type=\"Float32\" Name=\"" << a.name
          << "\" NumberOfComponents=\"1\" format=\"appended\" offset=\"" << a.offset << "\"/>\n";
    v << "      </PointData>\n";
    v << "      <CellData/>\n";
    v << "    </Piece>\n";
    v << "  </ImageData>\n";
    v << "  <AppendedData encoding=\"raw\">\n";
    v << '_';
    std::vector<char> buf(1 << 20);
    for (const Arr &a : arrays) {
        const uint32_t nb = uint32_t(node_bytes);
        v.write(reinterpret_cast<const char *>(&nb), sizeof(nb));
        std::ifstream in(a.path, std::ios::binary);
        while (in) {
            in.read(buf.data(), std::streamsize(buf.size()));
            const std::streamsize got = in.gcount();
            if (got > 0) v.write(buf.data(), got);
        }
    }
    v << "\n  </AppendedData>\n";
    v << "</VTKFile>\n";
    return v ? SSMD_OK : SSMD_EINVAL;
}

int ssmd_checkpoint_write(const ssmd_params *p, const ssmd_fields *d, uint64_t step, const char *path, void *stream) {
    if (!p || !d || !path) return SSMD_EINVAL;
    if (ssmd_validate(p) != 0) return SSMD_EINVAL;
    const std::vector<Buf> bufs = state_buffers(*p, *d);
    for (const Buf &b : bufs)
        if (!b.ptr) return SSMD_ENOTREADY;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (const hipError_t e = hipStreamSynchronize(s); e != hipSuccess) return int(e);
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return SSMD_EINVAL;
    CkptHeader h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.abi = SSMD_ABI_VERSION;
    h.params_bytes = int32_t(sizeof(ssmd_params));
    h.step = step;
    h.p = *p;
    f.write(reinterpret_cast<const char *>(&h), sizeof(h));
    std::vector<char> host(kChunk);
    for (const Buf &b : bufs)
        for (size_t off = 0; off < b.bytes; off += kChunk) {
            const size_t len = std::min(kChunk, b.bytes - off);
            hipError_t e = hipMemcpyAsync(host.data(), static_cast<const char *>(b.ptr) + off, len,
                                          hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return int(e);
            f.write(host.data(), std::streamsize(len));
        }
    return f ? SSMD_OK : SSMD_EINVAL;
}

int ssmd_checkpoint_read(const ssmd_params *p, ssmd_fields *d, uint64_t *step, const char *path, void *stream) {
    if (!p || !d || !step || !path) return SSMD_EINVAL;
    if (ssmd_validate(p) != 0) return SSMD_EINVAL;
    std::ifstream f(path, std::ios::binary);
    if (!f) return SSMD_EINVAL;
    CkptHeader h{};
    f.read(reinterpret_cast<char *>(&h), sizeof(h));
    if (!f || std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0 || h.abi != SSMD_ABI_VERSION ||
        h.params_bytes != int32_t(sizeof(ssmd_params)) || !same_case(h.p, *p))
        return SSMD_EINVAL;
    const std::vector<Buf> bufs = state_buffers(*p, *d);
    for (const Buf &b : bufs)
        if (!b.ptr) return SSMD_ENOTREADY;
    // Ghost-cell slots that no kernel writes keep their t = 0 values in both
    // buffers of each pair for the whole run: initialize, then overwrite the
    // state the file carries.
    if (const int rc = ssmd_init(p, d, stream); rc != 0) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (const hipError_t e = hipStreamSynchronize(s); e != hipSuccess) return int(e);
    std::vector<char> host(kChunk);
    for (const Buf &b : bufs)
        for (size_t off = 0; off < b.bytes; off += kChunk) {
            const size_t len = std::min(kChunk, b.bytes - off);
            f.read(host.data(), std::streamsize(len));
            if (!f) return SSMD_EINVAL;
            hipError_t e = hipMemcpyAsync(static_cast<char *>(b.ptr) + off, host.data(), len, hipMemcpyHostToDevice,
                                          s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return int(e);
        }
    *step = h.step;
    return SSMD_OK;
}

int ssmd_write_info(const ssmd_params *p, const char *dir, const char *id, const char *velocity_set, uint64_t nsteps,
                    uint64_t save_every) {
    if (!p || !dir || !id || !velocity_set) return SSMD_EINVAL;
    const std::filesystem::path path = std::filesystem::path(dir) / (std::string(id) + "_info.txt");
    std::ofstream f(path, std::ios::out | std::ios::trunc);
    if (!f) return SSMD_EINVAL;
    // The reference prints its compile-time Reynolds and Weber constants; here
    // they follow from the runtime parameters by inverting the formulas that
    // set the viscosities and sigma: nu = u D / Re (globalFunctions.cuh:234-242)
    // and sigma = u_oil^2 D_oil / We (constants.cuh:91).  A disabled jet prints 0.
    auto reynolds = [](const ssmd_jet &j, double nu) {
        return j.enabled ? dimensionless(j.u * jet_diameter(j) / nu) : std::string("0");
    };
    const std::string weber =
        p->oil.enabled && p->sigma > 0.0 ? dimensionless(p->oil.u * p->oil.u * jet_diameter(p->oil) / p->sigma)
                                         : std::string("0");
    auto diam = [](const ssmd_jet &j) { return j.enabled ? static_cast<long long>(jet_diameter(j)) : 0LL; };
    f << "---------------------------- SIMULATION METADATA ----------------------------\n"
      << "ID:                 " << id << '\n'
      << "Velocity set:       " << velocity_set << '\n'
      << "Water velocity:     " << (p->water.enabled ? float(p->water.u) : 0.0f) << '\n'
      << "Oil velocity:       " << (p->oil.enabled ? float(p->oil.u) : 0.0f) << '\n'
      << "Water Reynolds:     " << reynolds(p->water, p->visc_water) << '\n'
      << "Oil Reynolds:       " << reynolds(p->oil, p->visc_oil) << '\n'
      << "Weber number:       " << weber << "\n\n"
      << "Domain size:        NX=" << p->nx << ", NY=" << p->ny << ", NZ=" << p->nz << '\n'
      << "Water diameter:     D=" << diam(p->water) << '\n'
      << "Oil diameter:       D=" << diam(p->oil) << '\n'
      << "Timesteps:          " << nsteps << '\n'
      << "Output interval:    " << save_every << "\n\n"
      << "-----------------------------------------------------------------------------\n";
    return f ? SSMD_OK : SSMD_EINVAL;
}

}  // extern "C"
