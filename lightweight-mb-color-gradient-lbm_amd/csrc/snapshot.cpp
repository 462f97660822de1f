// Restartable snapshots of a library-owned domain: its configuration, storage format, step count
// and full stored state (11 fields over the stored extent, the jet/RTI ghost planes included).
// The reference has no checkpoint (its main loop writes VTI output only, main.cu:38-55; the VTI
// holds 5 of the 11 fields and cannot restart a run); a snapshot restores a domain so that
// stepping on from it gives bitwise the results of the uninterrupted run.
//
// File (little-endian, this library's ABI version): a fixed header, the cglbm_config struct as
// stored in memory, then 11 * nx * nyt * nz binary32 values in cglbm_moments field order, compact
// x-fastest layout (cglbm_copy_out's). FP16-deviation states are written decoded: the decode is
// exact and encoding a decoded value gives back its binary16 bits, so the state round-trips.
// A 64-bit FNV-1a checksum of the data guards against truncation and corruption.
#include "cglbm.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "domain_internal.hpp"

namespace cglbm {
int set_error(int code, const char *msg);
}
using cglbm::set_error;

namespace {

constexpr char kMagic[8] = {'C', 'G', 'L', 'B', 'M', 'S', 'N', 'P'};
constexpr uint32_t kVersion = 1;

struct SnapshotHeader {
    char magic[8];
    uint32_t version;
    uint32_t config_bytes;  // sizeof(cglbm_config) of the writer
    int32_t abi;            // cglbm_abi_version() of the writer
    int32_t storage;        // CGLBM_STORE_*
    int64_t steps;          // steps the state has taken since it was initialised
    uint64_t nfloats;       // 11 * nx * nyt * nz
    uint64_t checksum;      // FNV-1a 64 of the nfloats * 4 data bytes
};

uint64_t fnv1a(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

uint64_t state_floats(const cglbm_config &c) {
    const int64_t nyt = c.ybc == CGLBM_Y_PERIODIC ? c.ny : c.ny + 2;
    return 11ull * static_cast<uint64_t>(c.nx) * static_cast<uint64_t>(nyt) * static_cast<uint64_t>(c.nz);
}

struct FileCloser {
    void operator()(FILE *f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

// Reads and checks the header and the configuration that follows it.
int read_head(FILE *f, SnapshotHeader &h, cglbm_config &cfg) {
    if (std::fread(&h, sizeof(h), 1, f) != 1) return set_error(CGLBM_EINVAL, "snapshot: truncated header");
    if (std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0) return set_error(CGLBM_EINVAL, "snapshot: not a cglbm snapshot");
    if (h.version != kVersion) return set_error(CGLBM_EINVAL, "snapshot: unknown format version");
    if (h.abi != cglbm_abi_version() || h.config_bytes != sizeof(cglbm_config))
        return set_error(CGLBM_EINVAL, "snapshot: written by another ABI version of the library");
    if (h.storage != CGLBM_STORE_F32 && h.storage != CGLBM_STORE_F16) return set_error(CGLBM_EINVAL, "snapshot: bad storage");
    if (std::fread(&cfg, sizeof(cfg), 1, f) != 1) return set_error(CGLBM_EINVAL, "snapshot: truncated configuration");
    if (cfg.nx < 1 || cfg.ny < 1 || cfg.nz < 1 || h.nfloats != state_floats(cfg) || h.steps < 0)
        return set_error(CGLBM_EINVAL, "snapshot: inconsistent header");
    return CGLBM_OK;
}

}  // namespace

extern "C" {

int cglbm_snapshot_write(cglbm_domain *dom, const char *path) {
    if (!dom || !path) return set_error(CGLBM_EINVAL, "null domain or path");
    cglbm_config cfg;
    if (int rc = cglbm_domain_config(dom, &cfg)) return rc;
    const uint64_t n = state_floats(cfg);
    std::vector<float> data;
    try {
        data.resize(n);
    } catch (...) {
        return set_error(CGLBM_ENOMEM, "snapshot: host buffer");
    }
    if (int rc = cglbm_copy_out(dom, 0x7FFu, data.data())) return rc;
    SnapshotHeader h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kVersion;
    h.config_bytes = sizeof(cglbm_config);
    h.abi = cglbm_abi_version();
    h.storage = dom->storage;
    h.steps = dom->steps;
    h.nfloats = n;
    h.checksum = fnv1a(data.data(), n * sizeof(float));
    File f(std::fopen(path, "wb"));
    if (!f) return set_error(CGLBM_EIO, "snapshot: cannot open the file for writing");
    if (std::fwrite(&h, sizeof(h), 1, f.get()) != 1 || std::fwrite(&cfg, sizeof(cfg), 1, f.get()) != 1 ||
        std::fwrite(data.data(), sizeof(float), n, f.get()) != n)
        return set_error(CGLBM_EIO, "snapshot: write failed");
    if (std::fclose(f.release()) != 0) return set_error(CGLBM_EIO, "snapshot: close failed");
    return CGLBM_OK;
}

int cglbm_snapshot_info(const char *path, cglbm_config *cfg, int32_t *storage, int64_t *steps) {
    if (!path) return set_error(CGLBM_EINVAL, "null path");
    File f(std::fopen(path, "rb"));
    if (!f) return set_error(CGLBM_EIO, "snapshot: cannot open the file");
    SnapshotHeader h;
    cglbm_config c;
    if (int rc = read_head(f.get(), h, c)) return rc;
    if (cfg) *cfg = c;
    if (storage) *storage = h.storage;
    if (steps) *steps = h.steps;
    return CGLBM_OK;
}

int cglbm_snapshot_read(cglbm_domain *dom, const char *path) {
    if (!dom || !path) return set_error(CGLBM_EINVAL, "null domain or path");
    File f(std::fopen(path, "rb"));
    if (!f) return set_error(CGLBM_EIO, "snapshot: cannot open the file");
    SnapshotHeader h;
    cglbm_config c;
    if (int rc = read_head(f.get(), h, c)) return rc;
    cglbm_config mine;
    if (int rc = cglbm_domain_config(dom, &mine)) return rc;
    if (std::memcmp(&c, &mine, sizeof(c)) != 0 || h.storage != dom->storage)
        return set_error(CGLBM_EINVAL, "snapshot: configuration or storage differs from the domain's");
    std::vector<float> data;
    try {
        data.resize(h.nfloats);
    } catch (...) {
        return set_error(CGLBM_ENOMEM, "snapshot: host buffer");
    }
    if (std::fread(data.data(), sizeof(float), h.nfloats, f.get()) != h.nfloats)
        return set_error(CGLBM_EINVAL, "snapshot: truncated data");
    if (fnv1a(data.data(), h.nfloats * sizeof(float)) != h.checksum)
        return set_error(CGLBM_EINVAL, "snapshot: checksum mismatch");
    if (int rc = cglbm_copy_in(dom, 0x7FFu, data.data())) return rc;
    dom->steps = h.steps;  // the inlet perturbation is keyed by the step index
    if (dom->batch) batch_member_changed(dom->batch);
    return CGLBM_OK;
}

}  // extern "C"
