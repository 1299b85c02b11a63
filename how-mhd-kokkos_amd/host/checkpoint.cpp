// Checkpoint files of the host shell (no counterpart in the reference): the
// state the time step needs to continue a run bit for bit, q(0) and the
// stage-0 faces, with the output step, step count and time of the loop.
#include "howmhd_host.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace {

constexpr char kMagic[8] = {'H', 'O', 'W', 'M', 'H', 'D', 'C', 'K'};
constexpr uint32_t kVersion = 1;

struct Header {
    char magic[8];
    uint32_t version;
    int32_t nx, ny, nz;
    int32_t out_step;
    int64_t total_steps;
    double time;
};

void fail(char* errbuf, int errlen, const std::string& msg) {
    if (errbuf && errlen > 0) std::snprintf(errbuf, (size_t)errlen, "%s", msg.c_str());
}

size_t cells(const howmhd_run* r) { return (size_t)(r->nx + 6) * (r->ny + 6) * (r->nz + 6); }

struct File {
    std::FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" int howmhd_host_write_checkpoint(const howmhd_run* r, const char* path, const double* q0, const double* b,
                                            int32_t out_step, int64_t total_steps, double time, char* errbuf,
                                            int errlen) {
    if (!r || !path || !q0 || !b) {
        fail(errbuf, errlen, "Error: write_checkpoint: null argument");
        return HOWMHD_EINVAL;
    }
    Header h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kVersion;
    h.nx = r->nx; h.ny = r->ny; h.nz = r->nz;
    h.out_step = out_step;
    h.total_steps = total_steps;
    h.time = time;
    const size_t n = cells(r);
    File f;
    f.f = std::fopen(path, "wb");
    bool ok = f.f != nullptr;
    ok = ok && std::fwrite(&h, sizeof(h), 1, f.f) == 1;
    ok = ok && std::fwrite(q0, sizeof(double), 8 * n, f.f) == 8 * n;
    ok = ok && std::fwrite(b, sizeof(double), 3 * n, f.f) == 3 * n;
    if (f.f) ok = (std::fclose(f.f) == 0) && ok;
    f.f = nullptr;
    if (!ok) {
        fail(errbuf, errlen, std::string("Error: cannot write checkpoint ") + path);
        return HOWMHD_EINVAL;
    }
    return HOWMHD_OK;
}

extern "C" int howmhd_host_read_checkpoint(const howmhd_run* r, const char* path, double* q0, double* b,
                                           int32_t* out_step, int64_t* total_steps, double* time, char* errbuf,
                                           int errlen) {
    if (!r || !path || !q0 || !b || !out_step || !total_steps || !time) {
        fail(errbuf, errlen, "Error: read_checkpoint: null argument");
        return HOWMHD_EINVAL;
    }
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) {
        fail(errbuf, errlen, std::string("Error: cannot open checkpoint ") + path);
        return HOWMHD_EINVAL;
    }
    Header h{};
    if (std::fread(&h, sizeof(h), 1, f.f) != 1 || std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0) {
        fail(errbuf, errlen, std::string("Error: not a checkpoint file: ") + path);
        return HOWMHD_EINVAL;
    }
    if (h.version != kVersion) {
        fail(errbuf, errlen, std::string("Error: unsupported checkpoint version ") + std::to_string(h.version));
        return HOWMHD_EINVAL;
    }
    if (h.nx != r->nx || h.ny != r->ny || h.nz != r->nz) {
        fail(errbuf, errlen,
             "Error: checkpoint grid " + std::to_string(h.nx) + " x " + std::to_string(h.ny) + " x " +
                 std::to_string(h.nz) + " does not match the input's " + std::to_string(r->nx) + " x " +
                 std::to_string(r->ny) + " x " + std::to_string(r->nz));
        return HOWMHD_EINVAL;
    }
    const size_t n = cells(r);
    if (std::fread(q0, sizeof(double), 8 * n, f.f) != 8 * n || std::fread(b, sizeof(double), 3 * n, f.f) != 3 * n) {
        fail(errbuf, errlen, std::string("Error: truncated checkpoint ") + path);
        return HOWMHD_EINVAL;
    }
    *out_step = h.out_step;
    *total_steps = h.total_steps;
    *time = h.time;
    return HOWMHD_OK;
}
