// sun_io.hip -- gauge-field checkpoints (host code).
//
// Two formats:
//   native     the lattice's device bytes exactly as stored (planes of 16-byte
//              records, sun_field.hpp), after an 80-byte little-endian header:
//                char magic[8] "SUNCKPT1"; u32 version (1); i32 nc; i32 dims[4];
//                i32 storage; i32 precision; f64 beta; u64 seed; u32 next_sweep;
//                u32 reserved (0); u64 payload_bytes; u32 payload_crc32; u32 reserved
//              Loading restores the field bit for bit; with the seed and the next
//              sweep number the Markov chain continues exactly as if never stopped
//              (the RNG has no other state).  CRC32 (IEEE 802.3) guards the payload.
//   reference  the layout of the reference's SaveBin_Gauge / ReadBin_Gauge
//              (io_gauge.cpp:429-583, :605-700): optionally a header
//                i32 grid[4]; beta as the lattice's Real (4 or 8 bytes);
//                u64 (size_t) bytes per saved real
//              then, for t, z, y, x (x fastest), the four full N x N complex link
//              matrices (mu = 0..3), row-major, as float or double.  Reading checks
//              the grid and beta against the lattice (io_gauge.cpp:585-593) and
//              projects the matrices onto the lattice's storage.
#include <hip/hip_runtime.h>

#include <array>
#include <new>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sun_internal.hpp"

using namespace sun;

namespace {

constexpr char kMagic[8] = {'S', 'U', 'N', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kVersion = 1;

#pragma pack(push, 1)
struct NativeHeader {
    char magic[8];
    uint32_t version;
    int32_t nc;
    int32_t dims[4];
    int32_t storage;
    int32_t precision;
    double beta;
    uint64_t seed;
    uint32_t next_sweep;
    uint32_t reserved0;
    uint64_t payload_bytes;
    uint32_t payload_crc32;
    uint32_t reserved1;
};
#pragma pack(pop)
static_assert(sizeof(NativeHeader) == 80, "native checkpoint header is 80 bytes");

// CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320), table driven
uint32_t crc32(const unsigned char* p, size_t n, uint32_t crc = 0) {
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[i] = c;
        }
        return t;
    }();
    crc = ~crc;
    for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

// canonical EO order [4][V][N][N][2] <-> the reference's lexicographic site order
// with the four directions of a site together: eo index of lexicographic site lex
size_t eo_of(const sun_lattice* L, size_t lex) {
    const int* X = L->d.dims;
    size_t r = lex;
    const int x0 = (int)(r % X[0]);
    r /= X[0];
    const int x1 = (int)(r % X[1]);
    r /= X[1];
    const int x2 = (int)(r % X[2]);
    const int x3 = (int)(r / X[2]);
    const int parity = (x0 + x1 + x2 + x3) & 1;
    return (lex >> 1) + (size_t)parity * (size_t)(L->g.half);
}

}  // namespace

// native checkpoint of lat appended to an open file (header + payload)
int sun::write_native(FILE* f, sun_lattice* L, uint64_t seed, uint32_t next_sweep, const char* path) {
    std::vector<unsigned char> payload(L->bytes);
    HIP_TRY(hipMemcpy(payload.data(), L->buf, L->bytes, hipMemcpyDeviceToHost));
    NativeHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.nc = L->d.nc;
    for (int i = 0; i < 4; i++) h.dims[i] = L->d.dims[i];
    h.storage = L->d.storage;
    h.precision = L->d.precision;
    h.beta = L->d.beta;
    h.seed = seed;
    h.next_sweep = next_sweep;
    h.payload_bytes = L->bytes;
    h.payload_crc32 = crc32(payload.data(), payload.size());
    if (fwrite(&h, sizeof(h), 1, f) != 1 || fwrite(payload.data(), 1, payload.size(), f) != payload.size())
        return fail(SUN_ERR_INVALID, "writing %s failed", path);
    return SUN_OK;
}

// native checkpoint of field bytes already on the host (a job not on the device)
bool sun::write_native_bytes(FILE* f, const sun_job_desc& d, const std::vector<unsigned char>& payload, uint64_t seed,
                             uint32_t next_sweep) {
    NativeHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.nc = d.nc;
    for (int i = 0; i < 4; i++) h.dims[i] = d.dims[i];
    h.storage = d.storage;
    h.precision = d.precision;
    h.beta = d.beta;
    h.seed = seed;
    h.next_sweep = next_sweep;
    h.payload_bytes = payload.size();
    h.payload_crc32 = crc32(payload.data(), payload.size());
    return fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(payload.data(), 1, payload.size(), f) == payload.size();
}

// reads a native checkpoint from an open file, checks it against the lattice
// description d and returns its payload (the field bytes) on the host
int sun::read_native(FILE* f, const sun_lattice_desc& d, size_t bytes, std::vector<unsigned char>& payload,
                     uint64_t* seed, uint32_t* next_sweep, const char* path) {
    NativeHeader h;
    if (fread(&h, sizeof(h), 1, f) != 1) return fail(SUN_ERR_INVALID, "%s: short header", path);
    if (memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion)
        return fail(SUN_ERR_INVALID, "%s: not a native checkpoint (magic/version)", path);
    if (h.nc != d.nc || h.storage != d.storage || h.precision != d.precision ||
        memcmp(h.dims, d.dims, sizeof(h.dims)) != 0)
        return fail(SUN_ERR_INVALID,
                    "%s: checkpoint of N=%d %dx%dx%dx%d storage %d precision %d does not match the lattice", path,
                    h.nc, h.dims[0], h.dims[1], h.dims[2], h.dims[3], h.storage, h.precision);
    if (h.payload_bytes != bytes) return fail(SUN_ERR_INVALID, "%s: payload size mismatch", path);
    try {
        payload.resize(bytes);
    } catch (const std::bad_alloc&) {
        return fail(SUN_ERR_NOMEM, "%s: host allocation of %zu bytes", path, bytes);
    }
    if (fread(payload.data(), 1, payload.size(), f) != payload.size())
        return fail(SUN_ERR_INVALID, "%s: truncated payload", path);
    if (crc32(payload.data(), payload.size()) != h.payload_crc32)
        return fail(SUN_ERR_INVALID, "%s: payload checksum mismatch", path);
    if (seed) *seed = h.seed;
    if (next_sweep) *next_sweep = h.next_sweep;
    return SUN_OK;
}

namespace {

int save_native(sun_lattice* L, const char* path, uint64_t seed, uint32_t next_sweep) {
    File fo;
    fo.f = fopen(path, "wb");
    if (!fo.f) return fail(SUN_ERR_INVALID, "cannot open %s for writing", path);
    SUN_TRY(write_native(fo.f, L, seed, next_sweep, path));
    const int rc = fclose(fo.f);
    fo.f = nullptr;
    return rc == 0 ? SUN_OK : fail(SUN_ERR_INVALID, "writing %s failed", path);
}

int load_native(sun_lattice* L, const char* path, uint64_t* seed, uint32_t* next_sweep) {
    File fi;
    fi.f = fopen(path, "rb");
    if (!fi.f) return fail(SUN_ERR_INVALID, "cannot open %s", path);
    std::vector<unsigned char> payload;
    SUN_TRY(read_native(fi.f, L->d, L->bytes, payload, seed, next_sweep, path));
    HIP_TRY(hipMemcpy(L->buf, payload.data(), L->bytes, hipMemcpyHostToDevice));
    return SUN_OK;
}

int save_reference(sun_lattice* L, const char* path, bool header, int save_prec) {
    const int n2 = L->d.nc * L->d.nc;
    const size_t V = (size_t)L->g.V, per_link = (size_t)n2 * 2;
    std::vector<double> canon(4 * V * per_link);
    SUN_TRY(sun_links_to_host(L, canon.data(), SUN_LAYOUT_CANONICAL));
    File fo;
    fo.f = fopen(path, "wb");
    if (!fo.f) return fail(SUN_ERR_INVALID, "cannot open %s for writing", path);
    bool ok = true;
    if (header) {
        int32_t grid[4] = {L->d.dims[0], L->d.dims[1], L->d.dims[2], L->d.dims[3]};
        ok &= fwrite(grid, sizeof(grid), 1, fo.f) == 1;
        if (L->d.precision == SUN_PREC_F32) {
            const float b = (float)L->d.beta;
            ok &= fwrite(&b, sizeof(b), 1, fo.f) == 1;
        } else {
            const double b = L->d.beta;
            ok &= fwrite(&b, sizeof(b), 1, fo.f) == 1;
        }
        const uint64_t confprec = (uint64_t)save_prec;
        ok &= fwrite(&confprec, sizeof(confprec), 1, fo.f) == 1;
    }
    std::vector<double> site(4 * per_link);
    std::vector<float> sitef(4 * per_link);
    for (size_t lex = 0; lex < V && ok; lex++) {
        const size_t e = eo_of(L, lex);
        for (int mu = 0; mu < 4; mu++)
            memcpy(&site[mu * per_link], &canon[((size_t)mu * V + e) * per_link], per_link * sizeof(double));
        if (save_prec == 8) {
            ok &= fwrite(site.data(), sizeof(double), site.size(), fo.f) == site.size();
        } else {
            for (size_t i = 0; i < site.size(); i++) sitef[i] = (float)site[i];
            ok &= fwrite(sitef.data(), sizeof(float), sitef.size(), fo.f) == sitef.size();
        }
    }
    if (fclose(fo.f) != 0) ok = false;
    fo.f = nullptr;
    return ok ? SUN_OK : fail(SUN_ERR_INVALID, "writing %s failed", path);
}

int load_reference(sun_lattice* L, const char* path, bool header, int save_prec) {
    const int n2 = L->d.nc * L->d.nc;
    const size_t V = (size_t)L->g.V, per_link = (size_t)n2 * 2;
    File fi;
    fi.f = fopen(path, "rb");
    if (!fi.f) return fail(SUN_ERR_INVALID, "cannot open %s", path);
    if (header) {
        // beta was written with the writing lattice's precision (4 or 8 bytes): the
        // file size tells which
        fseek(fi.f, 0, SEEK_END);
        const long size = ftell(fi.f);
        fseek(fi.f, 0, SEEK_SET);
        int32_t grid[4];
        if (fread(grid, sizeof(grid), 1, fi.f) != 1) return fail(SUN_ERR_INVALID, "%s: short header", path);
        for (int i = 0; i < 4; i++)
            if (grid[i] != L->d.dims[i])
                return fail(SUN_ERR_INVALID, "%s: grid %dx%dx%dx%d does not match the lattice", path, grid[0],
                            grid[1], grid[2], grid[3]);
        int found = 0;
        double beta = 0.0;
        uint64_t confprec = 0;
        unsigned char b[8] = {0};
        for (int bsize = 4; bsize <= 8 && !found; bsize += 4) {
            fseek(fi.f, 16, SEEK_SET);
            uint64_t cp = 0;
            if (fread(b, 1, bsize, fi.f) != (size_t)bsize || fread(&cp, sizeof(cp), 1, fi.f) != 1) continue;
            if ((cp != 4 && cp != 8) || size != (long)(16 + bsize + 8 + 4 * V * per_link * cp)) continue;
            if (bsize == 4) {
                float bf;
                memcpy(&bf, b, 4);
                beta = bf;
            } else {
                memcpy(&beta, b, 8);
            }
            confprec = cp;
            found = bsize;
        }
        if (!found) return fail(SUN_ERR_INVALID, "%s: header/size do not describe a %d-colour lattice", path, L->d.nc);
        // A 4-byte beta is (float)beta as written here, or -- from the reference's
        // SaveBin_Gauge with Real = float, which writes sizeof(float) bytes starting at
        // the double PARAMS::Beta (io_gauge.cpp:457) -- the double's first (low) 4 bytes
        bool beta_ok = found == 4 ? (float)beta == (float)L->d.beta : beta == L->d.beta;
        if (found == 4 && !beta_ok) {
            const double lb = L->d.beta;
            beta_ok = memcmp(b, &lb, 4) == 0;
            if (beta_ok) beta = lb;
        }
        if (!beta_ok) return fail(SUN_ERR_INVALID, "%s: beta %.17g does not match the lattice's %.17g", path, beta,
                                  L->d.beta);
        save_prec = (int)confprec;
        fseek(fi.f, 16 + found + 8, SEEK_SET);
    }
    std::vector<double> canon(4 * V * per_link);
    std::vector<double> site(4 * per_link);
    std::vector<float> sitef(4 * per_link);
    for (size_t lex = 0; lex < V; lex++) {
        if (save_prec == 8) {
            if (fread(site.data(), sizeof(double), site.size(), fi.f) != site.size())
                return fail(SUN_ERR_INVALID, "%s: truncated", path);
        } else {
            if (fread(sitef.data(), sizeof(float), sitef.size(), fi.f) != sitef.size())
                return fail(SUN_ERR_INVALID, "%s: truncated", path);
            for (size_t i = 0; i < site.size(); i++) site[i] = sitef[i];
        }
        const size_t e = eo_of(L, lex);
        for (int mu = 0; mu < 4; mu++)
            memcpy(&canon[((size_t)mu * V + e) * per_link], &site[mu * per_link], per_link * sizeof(double));
    }
    return sun_links_from_host(L, canon.data(), SUN_LAYOUT_CANONICAL);
}

}  // namespace

static int sun_lattice_save_body(sun_lattice* lat, const char* path, int format, int save_prec, uint64_t seed,
                                 uint32_t next_sweep) {
    if (!lat || !path) return fail(SUN_ERR_INVALID, "sun_lattice_save: NULL argument");
    HIP_TRY(hipSetDevice(lat->ctx->device));
    HIP_TRY(hipDeviceSynchronize());  // work queued on any stream completes first
    switch (format) {
        case SUN_FILE_NATIVE:
            return save_native(lat, path, seed, next_sweep);
        case SUN_FILE_REFERENCE:
        case SUN_FILE_REFERENCE_NOHEADER:
            if (save_prec != 4 && save_prec != 8)
                return fail(SUN_ERR_INVALID, "sun_lattice_save: save_prec %d (4 or 8)", save_prec);
            return save_reference(lat, path, format == SUN_FILE_REFERENCE, save_prec);
        default:
            return fail(SUN_ERR_INVALID, "sun_lattice_save: format %d", format);
    }
}

extern "C" int sun_lattice_save(sun_lattice* lat, const char* path, int format, int save_prec, uint64_t seed,
                                uint32_t next_sweep) {
    return guard_alloc("sun_lattice_save", [&] {
        return sun_lattice_save_body(lat, path, format, save_prec, seed, next_sweep);
    });
}

static int sun_lattice_load_body(sun_lattice* lat, const char* path, int format, int save_prec, uint64_t* seed,
                                 uint32_t* next_sweep) {
    if (!lat || !path) return fail(SUN_ERR_INVALID, "sun_lattice_load: NULL argument");
    HIP_TRY(hipSetDevice(lat->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    switch (format) {
        case SUN_FILE_NATIVE:
            return load_native(lat, path, seed, next_sweep);
        case SUN_FILE_REFERENCE:
        case SUN_FILE_REFERENCE_NOHEADER:
            if (format == SUN_FILE_REFERENCE_NOHEADER && save_prec != 4 && save_prec != 8)
                return fail(SUN_ERR_INVALID, "sun_lattice_load: save_prec %d (4 or 8)", save_prec);
            return load_reference(lat, path, format == SUN_FILE_REFERENCE, save_prec);
        default:
            return fail(SUN_ERR_INVALID, "sun_lattice_load: format %d", format);
    }
}

extern "C" int sun_lattice_load(sun_lattice* lat, const char* path, int format, int save_prec, uint64_t* seed,
                                uint32_t* next_sweep) {
    return guard_alloc("sun_lattice_load", [&] {
        return sun_lattice_load_body(lat, path, format, save_prec, seed, next_sweep);
    });
}
